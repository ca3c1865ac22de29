"""Pins of the oracle's restatement of the coupled coarse/fine md:: HMC
(field-multi-md.h, with the field-multi.h primitives it calls).

None of these checks retypes the oracle's formulas. They pin it by:
  * finite differences: the fine force is minus the derivative of
    S_gauge + S_xi / (1 + md_alpha) for every stout_type class (the reference
    applies no (1 + md_alpha) to the stout term), the coarse force is
    -i alpha/(1+alpha) dGb/dt Gb^+ along the fine momenta, and get_DQ is dQ/dt;
  * gauge covariance of Gb and invariance of S_xi (a wrong staple path breaks it);
  * closed forms on unit links;
  * reversibility of zeta_T / zeta_H and Delta H -> 0 as dt -> 0: the fine
    force carries 1/(1+alpha) of the xi force and the coarse links absorb the
    rest, so H = K + S + S_xi is conserved by the continuous flow only if the
    fine force, the coarse velocity and their factors are all right;
  * the heat bath's one-link distribution against the SU(3) Weyl integral
    of exp(k Re Tr W).
The fine lattice of the derivative and covariance checks has coarse extents
3, 3, 3, 4, so no shift by a coarse site aliases another.
"""
import numpy as np
import pytest
import scipy.linalg as sl

LF = (6, 6, 6, 8)     # coarse (3, 3, 3, 4)
LS = (4, 4, 4, 4)     # coarse (2, 2, 2, 2): dynamics checks


def _vol(L):
    return int(np.prod(L))


def _site(x, L):
    return x[0] + L[0] * (x[1] + L[1] * (x[2] + L[2] * x[3]))


def _coarse(L):
    return tuple(v // 2 for v in L)


def _setup(orc, L, seed, rho=0.15, xi0=1.0, alpha=0.5):
    U = orc.haar_links(L, seed)
    Uc = orc.haar_links(_coarse(L), seed + 1)
    rng = np.random.default_rng(seed)
    xi = rng.uniform(0.5, 1.5, size=(_vol(_coarse(L)), 4))
    return U, Uc, xi, orc.multi(rho, xi0, alpha)


def _rotate(orc, U, site, mu, a, eps):
    T = orc.generators()
    Ucx = orc.as_cplx(U).copy()
    Ucx[site, mu] = sl.expm(1j * eps * T[a]) @ Ucx[site, mu]
    return orc.as_real(Ucx)


def _stout_type(x, mu):
    odd = [i for i in range(4) if x[i] % 2]
    if not odd:
        return 1
    if len(odd) == 1:
        return 3 if odd[0] == mu else 2
    if len(odd) == 2 and mu in odd:
        return 4
    return 0


# fine links of every stout_type class (field-multi.h:384-407), several nu
LINKS = [((0, 0, 0, 0), 0, 0), ((2, 4, 0, 6), 3, 5), ((4, 2, 2, 0), 1, 2),
         ((1, 0, 2, 0), 0, 3), ((0, 0, 4, 7), 3, 7),
         ((0, 1, 0, 0), 0, 1), ((2, 0, 0, 3), 1, 6), ((4, 0, 5, 2), 3, 4),
         ((1, 1, 0, 0), 0, 2), ((0, 3, 2, 5), 3, 0), ((4, 5, 1, 2), 2, 6),
         ((1, 1, 1, 0), 3, 4), ((0, 1, 3, 0), 0, 7)]


def test_links_cover_every_type():
    assert sorted({_stout_type(x, mu) for x, mu, _ in LINKS}) == [0, 1, 2, 3, 4]


@pytest.mark.parametrize("act_kw", [dict(type=0, beta=5.7), dict(type=1, beta=2.6, c1=-0.331)])
def test_fine_force_is_minus_derivative(orc, act_kw):
    """md::get_Fforce = -d/d omega_a [S_gauge + S_xi/(1+alpha)] (4th-order differences)."""
    act = orc.action(**act_kw)
    U, Uc, xi, mp = _setup(orc, LF, 301)
    F = orc.as_cplx(orc.multi_fforce(LF, U, Uc, xi, act, mp))
    T = orc.generators()

    def S(V):
        return orc.gauge_action(LF, V, act) + orc.xi_energy(LF, V, Uc, xi, mp) / (1. + mp.md_alpha)

    h = 2e-3
    for x, mu, a in LINKS:
        s = _site(x, LF)
        v = [S(_rotate(orc, U, s, mu, a, e * h)) for e in (2, 1, -1, -2)]
        fd = -(-v[0] + 8 * v[1] - 8 * v[2] + v[3]) / (12 * h)
        fa = np.trace(T[a] @ F[s, mu]).real
        assert abs(fd - fa) < 2e-7 * max(1.0, abs(fa)), (x, mu, a, fd, fa)


def test_xi_part_is_nonzero_on_every_type(orc):
    """The stout term is exercised: removing S_xi changes the force on the
    links of types 1-4 and leaves type 0 unchanged."""
    act = orc.action(0, 5.7)
    U, Uc, xi, mp = _setup(orc, LF, 302)
    F = orc.as_cplx(orc.multi_fforce(LF, U, Uc, xi, act, mp))
    F0 = orc.as_cplx(orc.multi_fforce(LF, U, Uc, np.zeros_like(xi), act, mp))
    Fg = orc.as_cplx(orc.get_force(LF, U, act))
    assert np.abs(F0 - Fg).max() < 1e-13
    for x, mu, _ in LINKS:
        s = _site(x, LF)
        d = np.abs(F[s, mu] - Fg[s, mu]).max()
        assert (d > 1e-4) if _stout_type(x, mu) else (d == 0.0), (x, mu, d)


def _flow(orc, L, U, P, t):
    V = U.copy()
    orc.evolve_gauge_field(L, V, P, t, 0)   # U -> e^{i t P} U
    return V


def _d4(f, h):
    return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)


def test_coarse_force_is_blocked_link_velocity(orc):
    """get_Cforce = -i alpha/(1+alpha) (dGb/dt) Gb^+ with U(t) = e^{itP} U."""
    U, _, _, mp = _setup(orc, LF, 311)
    P = orc.init_momentum(LF, 312, 0)
    CF = orc.as_cplx(orc.multi_cforce(LF, U, P, mp))
    Gb = orc.as_cplx(orc.multi_gb(LF, U, mp))
    dGb = _d4(lambda t: orc.as_cplx(orc.multi_gb(LF, _flow(orc, LF, U, P, t), mp)), 1e-3)
    k = mp.md_alpha / (1. + mp.md_alpha)
    ref = -1j * k * dGb @ np.conj(np.swapaxes(Gb, -1, -2))
    assert np.abs(CF - ref).max() < 1e-8 * np.abs(ref).max()
    # Hermitian traceless: Gb stays in SU(3) along the flow
    assert np.abs(CF - np.conj(np.swapaxes(CF, -1, -2))).max() < 1e-13
    assert np.abs(np.trace(CF, axis1=-2, axis2=-1)).max() < 1e-13


def test_DQ_is_Q_velocity(orc):
    U, _, _, mp = _setup(orc, LF, 321)
    P = orc.init_momentum(LF, 322, 0)
    DQ = orc.as_cplx(orc.multi_DQ(LF, U, P, mp))
    ref = _d4(lambda t: orc.as_cplx(orc.multi_Q(LF, _flow(orc, LF, U, P, t), mp)), 1e-3)
    assert np.abs(DQ - ref).max() < 1e-8 * np.abs(ref).max()


def _gauge(orc, L, U, g):
    X = orc.as_cplx(U).copy().reshape(L[3], L[2], L[1], L[0], 4, 3, 3)
    G = g.reshape(L[3], L[2], L[1], L[0], 3, 3)
    for mu in range(4):
        ax = 3 - mu
        Gs = np.roll(G, -1, axis=ax)
        X[..., mu, :, :] = G @ X[..., mu, :, :] @ np.conj(np.swapaxes(Gs, -1, -2))
    return orc.as_real(X.reshape(-1, 4, 3, 3))


def test_gauge_covariance(orc):
    """Gb -> g(2xc) Gb g(2xc + 2mu)^+ under a fine gauge transformation, and
    S_xi is invariant when the coarse links transform with g at even sites."""
    U, Uc, xi, mp = _setup(orc, LF, 331)
    LC = _coarse(LF)
    g = orc.as_cplx(orc.haar_links(LF, 332))[:, 0]
    Ug = _gauge(orc, LF, U, g)
    ge = g.reshape(LF[3], LF[2], LF[1], LF[0], 3, 3)[::2, ::2, ::2, ::2].reshape(-1, 3, 3)
    Ucg = _gauge(orc, LC, Uc, ge)
    Gb = orc.as_cplx(orc.multi_gb(LF, U, mp))
    Gbg = orc.as_cplx(orc.multi_gb(LF, Ug, mp))
    Gbt = orc.as_cplx(_gauge(orc, LC, orc.as_real(Gb), ge))
    assert np.abs(Gbg - Gbt).max() < 1e-12
    s0 = orc.xi_energy(LF, U, Uc, xi, mp)
    assert abs(orc.xi_energy(LF, Ug, Ucg, xi, mp) - s0) < 1e-11 * abs(s0)


def test_unit_links_closed_forms(orc):
    L = LF
    U = orc.unit_links(L)
    Uc = orc.unit_links(_coarse(L))
    xi = np.full((_vol(_coarse(L)), 4), 0.7)
    mp = orc.multi(0.1, 0.7, 0.5)
    # Q = 0, Gb = 1, Re Tr(1) = 3 per coarse link
    assert np.abs(orc.as_cplx(orc.multi_gb(L, U, mp)) - np.eye(3)).max() < 1e-15
    assert abs(orc.xi_energy(L, U, Uc, xi, mp) - (-(1.5) * 3 * 0.7 * 4 * _vol(_coarse(L)))) < 1e-9
    act = orc.action(0, 6.0)
    h, parts = orc.multi_hamiltonian(L, U, np.zeros_like(U), Uc, xi, act, mp)
    assert parts[0] == 0.0 and abs(parts[1] + 6.0 / 3 * 18 * _vol(L)) < 1e-8
    assert abs(h - parts.sum()) < 1e-9


def _multi_state(orc, L, seed, eps=0.25):
    """Fine links near a smooth configuration (not Haar: forces of thermal
    size), coarse links near the blocked links, momenta of unit variance."""
    U = orc.unit_links(L)
    orc.evolve_gauge_field(L, U, orc.init_momentum(L, seed, 0), eps, 0)
    mp = orc.multi(0.1, 1.0, 0.5)
    Uc = orc.multi_gb(L, U, mp)
    orc.evolve_gauge_field(_coarse(L), Uc, orc.init_momentum(_coarse(L), seed + 1, 0), 0.3, 0)
    xi = np.full((_vol(_coarse(L)), 4), 1.0)
    return U, Uc, xi, mp


def test_zeta_T_reversible(orc):
    L = LS
    U0, C0, _, mp = _multi_state(orc, L, 341)
    P = orc.init_momentum(L, 342, 0)
    U, C = U0.copy(), C0.copy()
    orc.multi_zeta_T(L, U, P, C, 0.2, 0, mp)
    assert np.abs(orc.as_cplx(C) - orc.as_cplx(C0)).max() > 1e-3
    orc.multi_zeta_T(L, U, -P, C, 0.2, 0, mp)
    assert np.abs(orc.as_cplx(U) - orc.as_cplx(U0)).max() < 1e-12
    assert np.abs(orc.as_cplx(C) - orc.as_cplx(C0)).max() < 1e-12


def _dH(orc, L, which, n, tau, U0, P0, C0, xi, act, mp):
    U, P, C = U0.copy(), P0.copy(), C0.copy()
    h0, _ = orc.multi_hamiltonian(L, U, P, C, xi, act, mp)
    orc.multi_integrate(L, U, P, C, xi, act, mp, which, n, tau / n)
    h1, _ = orc.multi_hamiltonian(L, U, P, C, xi, act, mp)
    return h1 - h0, U, P, C


def test_zeta_H_reversible_and_conserves_H(orc):
    L = LS
    act = orc.action(0, 5.7)
    U0, C0, xi, mp = _multi_state(orc, L, 351)
    P0 = orc.init_momentum(L, 352, 0)
    dh, U, P, C = _dH(orc, L, orc.MULTI_ZETA_H, 4, 0.4, U0, P0, C0, xi, act, mp)
    orc.multi_integrate(L, U, -P, C, xi, act, mp, orc.MULTI_ZETA_H, 4, 0.1)
    assert np.abs(orc.as_cplx(U) - orc.as_cplx(U0)).max() < 1e-10
    assert np.abs(orc.as_cplx(C) - orc.as_cplx(C0)).max() < 1e-10
    # Delta H -> 0 with dt: the coupled flow conserves K + S + S_xi
    d = [abs(_dH(orc, L, orc.MULTI_ZETA_H, n, 0.4, U0, P0, C0, xi, act, mp)[0]) for n in (2, 4, 8)]
    assert d[0] / d[1] > 3.0 and d[1] / d[2] > 3.0, d
    assert d[2] < 1e-3, d


def test_multi_fg_keeps_coarse_and_scales_dt4(orc):
    """md::force_gradient_integrator moves the fine fields only; with the
    coarse links fixed, S_xi is an ordinary potential for the fine links but
    the fine force carries only 1/(1+alpha) of it, so H' = K + S + S_xi/(1+alpha)
    is the conserved quantity, to O(dt^4)."""
    L = LS
    act = orc.action(0, 5.7)
    U0, C0, xi, mp = _multi_state(orc, L, 361)
    P0 = orc.init_momentum(L, 362, 0)

    def hprime(U, P):
        h, parts = orc.multi_hamiltonian(L, U, P, C0, xi, act, mp)
        return parts[0] + parts[1] + parts[2] / (1. + mp.md_alpha)

    d = []
    for n in (3, 6):
        U, P, C = U0.copy(), P0.copy(), C0.copy()
        h0 = hprime(U, P)
        orc.multi_integrate(L, U, P, C, xi, act, mp, orc.MULTI_FG, n, 0.6 / n)
        assert np.abs(C - C0).max() < 1e-15
        d.append(abs(hprime(U, P) - h0))
    assert 16 * 0.5 < d[0] / d[1] < 16 * 2.0, d


def _weyl_mean_retr(k, n=600):
    """<Re Tr W> for dP ~ exp(k Re Tr W) dHaar on SU(3), by the Weyl
    integration formula over eigenphases (t1, t2, -t1-t2)."""
    t = (np.arange(n) + 0.5) * 2 * np.pi / n - np.pi
    t1, t2 = np.meshgrid(t, t, indexing="ij")
    t3 = -t1 - t2
    z = [np.exp(1j * t1), np.exp(1j * t2), np.exp(1j * t3)]
    vdm = np.abs(z[0] - z[1]) ** 2 * np.abs(z[0] - z[2]) ** 2 * np.abs(z[1] - z[2]) ** 2
    r = np.cos(t1) + np.cos(t2) + np.cos(t3)
    w = vdm * np.exp(k * r)
    return (w * r).sum() / w.sum()


def test_weyl_integral_reference_values():
    # k = 0: Haar, <Tr W> = 0; small k: <Re Tr W> = k/2 + O(k^2) (<|Tr W|^2> = 1)
    assert abs(_weyl_mean_retr(0.0)) < 1e-12
    assert abs(_weyl_mean_retr(1e-3) - 0.5e-3) < 1e-6


@pytest.mark.slow
def test_heatbath_one_link_distribution(orc):
    """simplest_metropolis samples exp((1+alpha) XI0 Re Tr(C Gb^+)): the mean
    of Re Tr(C Gb^+) over the coarse links matches the Weyl integral."""
    L = LS
    U = orc.haar_links(L, 371)
    mp = orc.multi(0.1, 1.0, 0.5)
    k = (1. + mp.md_alpha) * mp.xi0
    Gb = orc.as_cplx(orc.multi_gb(L, U, mp))
    C = orc.haar_links(_coarse(L), 372)
    vals = []
    for t in range(1500):
        orc.multi_heatbath(L, C, U, mp, 373, t)
        if t >= 300 and t % 5 == 0:
            W = orc.as_cplx(C) @ np.conj(np.swapaxes(Gb, -1, -2))
            vals.append(np.trace(W, axis1=-2, axis2=-1).real.mean())
    ref = _weyl_mean_retr(k)
    assert abs(np.mean(vals) - ref) < 0.06, (np.mean(vals), ref)


def test_run_hmc_multi_bookkeeping(orc):
    L = LS
    act = orc.action(0, 5.7)
    U, Uc, _, mp = _multi_state(orc, L, 381)
    h = orc.hmc(trajectory_length=3, dt=0.1, num_forced_accept=1, seed=382)
    U1, C1 = U.copy(), Uc.copy()
    nacc, res = orc.run_hmc_multi(L, U1, C1, act, h, mp, 0, 3)
    assert res[0].accept == 1
    for r in res:
        assert abs(r.dh - (r.h_new - r.h_old)) < 1e-9
        assert abs(r.h_old - sum(r.parts_old)) < 1e-8 * abs(r.h_old)
        assert abs(r.h_new - sum(r.parts_new)) < 1e-8 * abs(r.h_new)
        assert r.accept == (1 if r.die_roll < np.exp(-r.dh) else 0) or r is res[0]
    assert nacc == sum(r.accept for r in res)
    assert abs(res[-1].plaq_fine - orc.avg_plaquette(L, U1)) < 1e-14
    assert abs(res[-1].plaq_coarse - orc.avg_plaquette(_coarse(L), C1)) < 1e-14
    # the same trajectories from the same state reproduce exactly
    U2, C2 = U.copy(), Uc.copy()
    orc.run_hmc_multi(L, U2, C2, act, h, mp, 0, 3)
    assert np.array_equal(U1, U2) and np.array_equal(C1, C2)
