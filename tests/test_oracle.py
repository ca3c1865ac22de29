"""Pins of the FP64 CPU oracle (oracle/oracle.cpp) against closed forms,
library routines (scipy), published known-answer vectors, an independent
numpy restatement of the lattice sums, finite differences and invariants.

The reference (hummingtree/mag) ships no tests, fixtures or golden vectors,
so every pin here comes from mathematics or from an independent code path.
Each test is chosen so that a plausible mistake (a dropped staple, a wrong
sign, a transposed product, a missing dagger) fails it.
"""
import numpy as np
import pytest
import scipy.linalg as sl

L4 = (4, 4, 4, 4)


def rand_herm_traceless(rng, scale, n=1):
    import pyoracle as O
    T = O.generators()
    q = rng.normal(size=(n, 8)) * scale
    return np.einsum("na,aij->nij", q, T)


# ------------------------------------------------------------------ RNG
def test_philox_known_answers(orc):
    # Random123 kat_vectors, philox4x32_10
    assert orc.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert orc.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert orc.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_momentum_moments(orc):
    """P = sum_a T_a g_a with g ~ N(0,1): <p_a> = 0, <p_a p_b> = delta_ab, <1/2 Re Tr P^2> = 4."""
    L = (8, 8, 8, 8)
    P = orc.as_cplx(orc.init_momentum(L, 1234, 0))
    T = orc.generators()
    p = np.einsum("aij,sxji->sxa", T, P).real.reshape(-1, 8)
    assert np.allclose(P, np.conj(np.swapaxes(P, -1, -2)), atol=1e-14)
    assert np.abs(np.trace(P, axis1=-2, axis2=-1)).max() < 1e-13
    n = p.shape[0]
    assert np.abs(p.mean(0)).max() < 5 / np.sqrt(n)
    cov = p.T @ p / n
    assert np.abs(cov - np.eye(8)).max() < 6 / np.sqrt(n)
    K = orc.kinetic(L, orc.init_momentum(L, 1234, 0)) / (4 * np.prod(L))
    assert abs(K - 4.0) < 6 * np.sqrt(8 / 2 / n)
    # different trajectories give different momenta; same inputs give the same
    assert not np.allclose(orc.init_momentum(L4, 1, 0), orc.init_momentum(L4, 1, 1))
    assert np.array_equal(orc.init_momentum(L4, 1, 3), orc.init_momentum(L4, 1, 3))


def test_haar_su3(orc):
    """Haar SU(3): unitary, det 1, <Tr U> = 0, <|Tr U|^2> = 1, <|Tr U|^4> = 2."""
    L = (8, 8, 8, 8)
    U = orc.as_cplx(orc.haar_links(L, 7)).reshape(-1, 3, 3)
    eye = np.eye(3)
    assert np.abs(U @ np.conj(np.swapaxes(U, -1, -2)) - eye).max() < 1e-13
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-13
    t = np.trace(U, axis1=-2, axis2=-1)
    n = len(t)
    assert abs(t.mean()) < 5 / np.sqrt(n)
    assert abs(np.mean(np.abs(t) ** 2) - 1) < 6 * np.sqrt(1 / n)
    assert abs(np.mean(np.abs(t) ** 4) - 2) < 6 * np.sqrt(20 / n)


# ------------------------------------------------------------------ algebra
def test_generators(orc):
    T = orc.generators()
    for a in range(8):
        assert np.allclose(T[a], np.conj(T[a].T))
        assert abs(np.trace(T[a])) < 1e-15
        for b in range(8):
            assert abs(np.trace(T[a] @ T[b]) - (a == b)) < 1e-15
    # lambda_2 / sqrt 2 and lambda_8 / sqrt 2 (Gell-Mann)
    assert np.allclose(T[1] * np.sqrt(2), [[0, -1j, 0], [1j, 0, 0], [0, 0, 0]])
    assert np.allclose(T[7] * np.sqrt(2), np.diag([1, 1, -2]) / np.sqrt(3))


@pytest.mark.parametrize("scale", [1e-6, 1e-3, 0.05, 0.3, 1.0, 2.5])
def test_expiQ_vs_scipy(orc, scale):
    rng = np.random.default_rng(int(scale * 1e6) + 1)
    Q = rand_herm_traceless(rng, scale, 200)
    E = orc.expiQ(Q)
    ref = np.array([sl.expm(1j * q) for q in Q])
    assert np.abs(E - ref).max() < 3e-15 * max(1, scale) * 3


def test_expiQ_closed_forms(orc):
    # diagonal Q: exp(i diag(t)) ; reflect branch (c0 < 0) via -Q ; Q = 0
    t = np.array([0.3, -0.1, -0.2])
    E = orc.expiQ(np.diag(t)[None])[0]
    assert np.abs(E - np.diag(np.exp(1j * t))).max() < 1e-15
    E2 = orc.expiQ(-np.diag(t)[None])[0]
    assert np.abs(E2 - np.diag(np.exp(-1j * t))).max() < 1e-15
    assert np.array_equal(orc.expiQ(np.zeros((1, 3, 3)))[0], np.eye(3))
    # degenerate spectrum (w -> 0 branch of xi0): diag(a, a, -2a)
    a = 0.2
    E3 = orc.expiQ(np.diag([a, a, -2 * a])[None])[0]
    assert np.abs(E3 - np.diag(np.exp(1j * np.array([a, a, -2 * a])))).max() < 1e-14


def _dexp_fd(Q, X, h=1e-5):
    """Central difference of expm(i(Q + eps X)) with a 4th-order stencil."""
    f = lambda e: sl.expm(1j * (Q + e * X))
    return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)


def test_DexpiQ_vs_finite_differences(orc):
    rng = np.random.default_rng(5)
    for scale in (0.1, 0.7, 1.5):
        Q = rand_herm_traceless(rng, scale, 6)
        X = rand_herm_traceless(rng, 1.0, 6)
        D = orc.DexpiQ(Q, X)
        for q, x, d in zip(Q, X, D):
            assert np.abs(d - _dexp_fd(q, x)).max() < 1e-9


def test_compute_Lambda_is_the_Q_derivative(orc):
    """Tr(dQ Lambda) = d/de Re Tr(Sigma' e^{i(Q + e dQ)} U) for Hermitian traceless dQ."""
    rng = np.random.default_rng(9)
    Q = rand_herm_traceless(rng, 0.5, 5)
    dQ = rand_herm_traceless(rng, 1.0, 5)
    S = rng.normal(size=(5, 3, 3)) + 1j * rng.normal(size=(5, 3, 3))
    U = orc.as_cplx(orc.haar_links(L4, 3)).reshape(-1, 3, 3)[:5]
    Lam = orc.compute_Lambda(Q, S, U)
    for q, x, s, u, lam in zip(Q, dQ, S, U, Lam):
        assert np.allclose(lam, np.conj(lam.T), atol=1e-13) and abs(np.trace(lam)) < 1e-13
        fd = np.trace(s @ _dexp_fd(q, x) @ u).real
        assert abs(np.trace(x @ lam).real - fd) < 1e-9 * max(1, abs(fd))


# ------------------------------------------------------------------ lattice
def _lat(U, L):
    """(V,4,18) -> complex (L3, L2, L1, L0, 4, 3, 3): axis of direction mu is 3 - mu."""
    import pyoracle as O
    return O.as_cplx(U).reshape(L[3], L[2], L[1], L[0], 4, 3, 3)


def _sh(A, mu, n=1):
    return np.roll(A, -n, axis=3 - mu)


def _dg(A):
    return np.conj(np.swapaxes(A, -1, -2))


def np_plaq_rect(U, L):
    """Independent numpy restatement: sum of Re Tr over plaquettes and 2x1/1x2 rectangles."""
    X = _lat(U, L)
    sp = sr = 0.0
    for mu in range(4):
        for nu in range(mu):
            Um, Un = X[..., mu, :, :], X[..., nu, :, :]
            P = Um @ _sh(Un, mu) @ _dg(_sh(Um, nu)) @ _dg(Un)
            sp += np.trace(P, axis1=-2, axis2=-1).real.sum()
            R1 = Um @ _sh(Um, mu) @ _sh(_sh(Un, mu), mu) @ _dg(_sh(_sh(Um, mu), nu)) @ _dg(_sh(Um, nu)) @ _dg(Un)
            R2 = Um @ _sh(Un, mu) @ _sh(_sh(Un, mu), nu) @ _dg(_sh(_sh(Um, nu), nu)) @ _dg(_sh(Un, nu)) @ _dg(Un)
            sr += np.trace(R1, axis1=-2, axis2=-1).real.sum() + np.trace(R2, axis1=-2, axis2=-1).real.sum()
    return sp, sr


def test_unit_gauge_values(orc):
    U = orc.unit_links(L4)
    V = np.prod(L4)
    assert orc.total_plaq(L4, U) == 18 * V
    assert orc.total_rectangular(L4, U) == 36 * V
    assert orc.avg_plaquette(L4, U) == 1.0
    assert orc.avg_wilson_loop(L4, U, (2, 1, 0), 2) == 3.0
    F = orc.get_force(L4, U, orc.action(1, 2.6, -0.331, n_stout=2))
    assert np.abs(F).max() < 1e-15


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (4, 6, 4, 8)])
def test_plaq_rect_vs_numpy(orc, L):
    U = orc.haar_links(L, 11)
    sp, sr = np_plaq_rect(U, L)
    assert abs(orc.total_plaq(L, U) - sp) < 1e-11 * max(1, abs(sp))
    assert abs(orc.total_rectangular(L, U) - sr) < 1e-11 * max(1, abs(sr))


def random_gauge(orc, L, seed):
    return orc.as_cplx(orc.haar_links(L, seed))[:, 0]  # one SU(3) per site


def gauge_transform(orc, U, g, L):
    """U_mu(x) -> g(x) U_mu(x) g(x+mu)^+"""
    Uc = orc.as_cplx(U).copy()
    G = g.reshape(L[3], L[2], L[1], L[0], 3, 3)
    X = Uc.reshape(L[3], L[2], L[1], L[0], 4, 3, 3)
    for mu in range(4):
        X[..., mu, :, :] = G @ X[..., mu, :, :] @ _dg(_sh(G, mu))
    return orc.as_real(X.reshape(-1, 4, 3, 3)), G


def test_gauge_invariance_and_force_covariance(orc):
    U = orc.haar_links(L4, 21)
    g = random_gauge(orc, L4, 22)
    Ug, G = gauge_transform(orc, U, g, L4)
    act = orc.action(1, 2.6, -0.331, n_stout=2, rho=0.1)
    assert abs(orc.gauge_action(L4, Ug, act) - orc.gauge_action(L4, U, act)) < 1e-11 * abs(orc.gauge_action(L4, U, act))
    assert abs(orc.avg_wilson_loop(L4, Ug, (1, 1, 0), 2) - orc.avg_wilson_loop(L4, U, (1, 1, 0), 2)) < 1e-13
    F = orc.as_cplx(orc.get_force(L4, U, act)).reshape(L4[3], L4[2], L4[1], L4[0], 4, 3, 3)
    Fg = orc.as_cplx(orc.get_force(L4, Ug, act)).reshape(F.shape)
    rot = G[..., None, :, :] @ F @ _dg(G)[..., None, :, :]
    assert np.abs(Fg - rot).max() < 1e-11


def _rotated_action(orc, U, act, site, mu, a, eps, L):
    T = orc.generators()
    Uc = orc.as_cplx(U).copy()
    x = [(site // int(np.prod(L[:i]))) % L[i] for i in range(4)]
    cls = orc.is_constrained(x, mu, act.mag)
    R = sl.expm(1j * eps * T[a])
    if cls == 0:
        Uc[site, mu] = R @ Uc[site, mu]
    else:
        # constrained classes 1 and 10 move the rotation between U(x) and U(x+mu)
        y = list(x)
        y[mu] = (y[mu] + 1) % L[mu]
        sy = y[0] + L[0] * (y[1] + L[1] * (y[2] + L[2] * y[3]))
        Uc[site, mu] = Uc[site, mu] @ np.conj(R.T)
        Uc[sy, mu] = R @ Uc[sy, mu]
    return orc.gauge_action(L, orc.as_real(Uc), act)


@pytest.mark.parametrize("act_kw", [
    dict(type=0, beta=6.0),
    dict(type=1, beta=2.6, c1=-0.331),
    dict(type=1, beta=1.04, c1=-1.4069),
    dict(type=1, beta=2.6, c1=-0.331, n_stout=1, rho=0.1),
    dict(type=1, beta=2.6, c1=-0.331, n_stout=3, rho=0.1),
    dict(type=0, beta=5.7, n_stout=2, rho=0.13),
    dict(type=1, beta=2.6, c1=-0.331, mag=2),
    dict(type=0, beta=6.0, mag=2, n_stout=1, rho=0.1),
])
def test_force_is_minus_action_derivative(orc, act_kw):
    """F_a(x,mu) = -dS/d omega_a for U -> e^{i omega T_a} U (4th-order central differences)."""
    act = orc.action(**act_kw)
    U = orc.haar_links(L4, 31)
    F = orc.as_cplx(orc.get_force(L4, U, act))
    T = orc.generators()
    h = 2e-3
    checked = 0
    for site, mu, a in [(0, 0, 0), (37, 1, 3), (130, 2, 7), (255, 3, 5), (2, 0, 2), (64, 1, 6)]:
        x = [(site // int(np.prod(L4[:i]))) % L4[i] for i in range(4)]
        if orc.is_constrained(x, mu, act.mag) == 100:
            assert np.abs(F[site, mu]).max() == 0.0
            continue
        S = [_rotated_action(orc, U, act, site, mu, a, e * h, L4) for e in (2, 1, -1, -2)]
        fd = -(-S[0] + 8 * S[1] - 8 * S[2] + S[3]) / (12 * h)
        fa = np.trace(T[a] @ F[site, mu]).real
        assert abs(fd - fa) < 2e-7 * max(1.0, abs(fa)), (site, mu, a, fd, fa)
        checked += 1
    assert checked >= 4


def test_force_matches_reference_derivative_identity(orc):
    """Tr(T_a F) = (beta/3) derivative_Iwasaki (field-hmc.h:702-712) = (beta/3) Re Tr(i T_a U Sigma^+)."""
    U = orc.haar_links(L4, 41)
    act = orc.action(1, 2.6, -0.331)
    F = orc.as_cplx(orc.get_force(L4, U, act))
    T = orc.generators()
    Uc = orc.as_cplx(U)
    c1 = -0.331
    c0 = 1 - 8 * c1
    for site, mu in [(5, 0), (77, 3)]:
        Sd = c0 * orc.staple_dagger(L4, U, site, mu) + c1 * orc.rectangular_dagger(L4, U, site, mu)
        for a in range(8):
            der = np.trace(T[a] @ Uc[site, mu] @ Sd * 1j).real
            assert abs(np.trace(T[a] @ F[site, mu]).real - 2.6 / 3 * der) < 1e-12


def test_staples_vs_path_products(orc):
    """Staple sums equal the sum of the generic path products they are built from."""
    U = orc.haar_links(L4, 51)
    x = [1, 2, 3, 0]
    site = x[0] + 4 * (x[1] + 4 * (x[2] + 4 * x[3]))
    mu = 2
    S = np.zeros((3, 3), complex)
    for nu in range(4):
        if nu == mu:
            continue
        S += orc.path_product(L4, U, x, [nu, mu, nu + 4]) + orc.path_product(L4, U, x, [nu + 4, mu, nu])
    assert np.abs(orc.staple_dagger(L4, U, site, mu) - np.conj(S.T)).max() < 1e-14
    # closed plaquette: Re Tr(U_mu(x) staple^+) summed over links = 4 * total_plaq
    tot = sum(np.trace(orc.as_cplx(U)[s, m] @ orc.staple_dagger(L4, U, s, m)).real
              for s in range(0, 256, 1) for m in range(4))
    assert abs(tot - 4 * orc.total_plaq(L4, U)) < 1e-10


def test_reunitarize(orc):
    U = orc.haar_links(L4, 61)
    V = U.copy()
    assert orc.reunitarize(L4, V) < 1e-13          # Haar links are SU(3) to rounding
    assert orc.reunitarize(L4, V) < 2e-15          # projection is idempotent
    rng = np.random.default_rng(0)
    W = U + 1e-3 * rng.normal(size=U.shape)
    dev = orc.reunitarize(L4, W)
    assert 1e-4 < dev < 1e-2
    M = orc.as_cplx(W).reshape(-1, 3, 3)
    assert np.abs(M @ _dg(M) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14
    assert np.abs(M[:, 2] - np.conj(np.cross(M[:, 0], M[:, 1]))).max() < 1e-15
    Z = np.zeros_like(U)
    assert orc.reunitarize(L4, Z) == 0.0 and not Z.any()   # all-zero links are skipped


def test_evolve_gauge_preserves_constrained_products(orc):
    """mag = 2: the product of the two fine links of a coarse link is unchanged by every class update."""
    U = orc.haar_links(L4, 71)
    P = orc.init_momentum(L4, 72, 0)
    def coarse_products(V):
        Vc = orc.as_cplx(V)
        out = []
        for x0 in (0, 2):
            for mu in range(4):
                x = [0, 0, 0, 0]
                x[0] = x0
                s = x[0] + 4 * (x[1] + 4 * (x[2] + 4 * x[3]))
                y = list(x)
                y[mu] += 1
                sy = y[0] + 4 * (y[1] + 4 * (y[2] + 4 * y[3]))
                out.append(Vc[s, mu] @ Vc[sy, mu])
        return np.array(out)
    before = coarse_products(U)
    orc.evolve_gauge_field(L4, U, P, 0.3, 2)
    assert np.abs(coarse_products(U) - before).max() < 1e-13
    assert np.abs(orc.as_cplx(U) - orc.as_cplx(orc.haar_links(L4, 71))).max() > 1e-2


def test_fg_reversibility_and_energy(orc):
    L = L4
    act = orc.action(1, 2.6, -0.331, n_stout=1, rho=0.1)
    U0 = orc.haar_links(L, 81)
    # mild start: a few smoothing steps bring forces to thermal size
    U = U0.copy()
    P = orc.init_momentum(L, 82, 0)
    h0, _ = orc.hamiltonian(L, U, P, act)
    orc.force_gradient_integrator(L, U, P, act, 4, 0.05, reunit=False)
    h1, _ = orc.hamiltonian(L, U, P, act)
    P *= -1
    orc.force_gradient_integrator(L, U, P, act, 4, 0.05, reunit=False)
    P *= -1
    assert np.abs(orc.as_cplx(U) - orc.as_cplx(U0)).max() < 1e-10
    assert np.abs(P - orc.init_momentum(L, 82, 0)).max() < 1e-10
    assert abs(h1 - h0) < 0.05 * abs(h0) ** 0.5


def _dH(orc, act, integ, steps, tau, U0, P0, L):
    U, P = U0.copy(), P0.copy()
    h0, _ = orc.hamiltonian(L, U, P, act)
    if integ == "fg":
        orc.force_gradient_integrator(L, U, P, act, steps, tau / steps, reunit=False)
    else:
        orc.leap_frog_integrator(L, U, P, act, steps, tau / steps)
    h1, _ = orc.hamiltonian(L, U, P, act)
    return h1 - h0


def test_dH_scaling_fg_dt4_leapfrog_dt2(orc):
    L = L4
    act = orc.action(0, 6.0)
    U0 = orc.haar_links(L, 91)
    # thermalise roughly with a few leapfrog trajectories (all accepted)
    for t in range(3):
        P = orc.init_momentum(L, 92, t)
        orc.leap_frog_integrator(L, U0, P, act, 20, 0.02)
    P0 = orc.init_momentum(L, 93, 0)
    tau = 0.5
    fg = [abs(_dH(orc, act, "fg", n, tau, U0, P0, L)) for n in (4, 8)]
    lf = [abs(_dH(orc, act, "lf", n, tau, U0, P0, L)) for n in (8, 16)]
    assert 2 ** 4 * 0.6 < fg[0] / fg[1] < 2 ** 4 * 1.6
    assert 2 ** 2 * 0.6 < lf[0] / lf[1] < 2 ** 2 * 1.6


def test_wilson_loop_relations(orc):
    U = orc.haar_links(L4, 101)
    # r = (1,0,0), T = 1: mean over the 3 permutations of the temporal plaquettes' Re Tr
    X = _lat(U, L4)
    tot = 0.0
    for i in range(3):
        Ui, Ut = X[..., i, :, :], X[..., 3, :, :]
        P = Ui @ _sh(Ut, i) @ _dg(_sh(Ui, 3)) @ _dg(Ut)
        tot += np.trace(P, axis1=-2, axis2=-1).real.sum()
    ref = tot / (3 * np.prod(L4))
    assert abs(orc.avg_wilson_loop(L4, U, (1, 0, 0), 1) - ref) < 1e-13
    # 2x2 loops: brute force through the generic path product at one site
    p = orc.path_product(L4, U, [0, 0, 0, 0], [0, 0, 3, 3, 4, 4, 7, 7])
    assert np.isfinite(np.trace(p).real)


def test_demon_energy_conservation(orc):
    """S_k(U) + E_k is conserved for the plaquette and rectangle sums; windows hold."""
    L = L4
    U = orc.haar_links(L, 111)
    act_p = (-orc.total_plaq(L, U), -orc.total_rectangular(L, U))
    E, (acc, rej) = orc.demon_microcanonical(L, U, 2, 40.0, 5)
    after = (-orc.total_plaq(L, U), -orc.total_rectangular(L, U))
    assert acc > 0 and rej > 0
    assert abs(after[0] + E[0] - act_p[0]) < 1e-9
    assert abs(after[1] + E[1] - act_p[1]) < 1e-9
    assert np.all(np.abs(E) < 40.0)


def test_run_hmc_bookkeeping(orc):
    L = L4
    U = orc.haar_links(L, 121)
    act = orc.action(0, 6.0)
    h = orc.hmc(trajectory_length=4, dt=0.1, num_forced_accept=2, seed=5)
    nacc, res, _ = orc.run_hmc(L, U, act, h, 0, 4)
    assert res[0].accept == 1 and res[1].accept == 1
    for i, r in enumerate(res):
        assert r.die_roll == orc.uniform(5, 3, i)
        if i >= 2:
            assert r.accept == int(r.die_roll < np.exp(r.h_old - r.h_new))
    assert abs(res[-1].avg_plaq - orc.avg_plaquette(L, U)) < 1e-15


# ------------------------------------------------------ mag-blocked lattices
def test_coarse_fine_identities(orc):
    """Pins of the coarse/fine tools by identities a mistake in a path, a
    direction or a dagger breaks: update_constrain makes the fine products
    equal the coarse links; the constrained plaquette of a fine field is the
    plaquette of its coarse links; an expanded lattice's constrained
    plaquette is the coarse plaquette; mag = 1 is the ordinary plaquette."""
    Lf, Lc = (8, 8, 4, 8), (4, 4, 2, 4)
    U = orc.haar_links(Lf, 3)
    Uc = orc.haar_links(Lc, 9)
    V = U.copy()
    orc.update_constrain(Lf, V, Uc, 2)
    assert np.abs(orc.coarsen(Lf, V, 2) - Uc).max() < 1e-13
    changed = np.abs(V - U).max(axis=2)
    origins = np.zeros((8, 4, 8, 8), bool)
    origins[::2, ::2, ::2, ::2] = True  # [x3, x2, x1, x0]
    assert not changed.reshape(8, 4, 8, 8, 4)[~origins].any()  # only the first link of each coarse link moves
    assert abs(orc.constrained_plaquette(Lf, U, 2) - orc.avg_plaquette(Lc, orc.coarsen(Lf, U, 2))) < 1e-14
    E = orc.expand(Lc, Uc, 2)
    assert abs(orc.constrained_plaquette(Lf, E, 2) - orc.avg_plaquette(Lc, Uc)) < 1e-14
    assert abs(orc.constrained_plaquette(Lf, U, 1) - orc.avg_plaquette(Lf, U)) < 1e-14
    unit = orc.unit_links(Lf)
    assert orc.avg_plaquette(Lf, E) != orc.avg_plaquette(Lc, Uc)  # the expansion is not the coarse lattice
    assert abs(orc.constrained_plaquette(Lf, unit, 2) - 1.0) < 1e-15
