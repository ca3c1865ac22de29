"""The coupled coarse/fine md:: HMC (field-multi-md.h) on the device, through
the C ABI, against the FP64 oracle: Gb, S_xi, the fine force of every
stout_type class, the coarse force, the heat bath (same Philox stream: the
same accept decisions), zeta_T, the three integrators, the Hamiltonian and
run_hmc_multi. At a larger lattice, properties that need no oracle: Delta H
of zeta_H falls with dt, and the heat bath samples the SU(3) one-link
distribution given by the Weyl integral.

Tolerances: 1e-12 relative for single evaluations (FP64 with a different
summation order from the oracle), 1e-9 after integrations of a few steps.
"""
import numpy as np
import pytest

import mag_amd as M
from mag_amd import ALG8, F64, LINK18, REAL1
from mag_amd import multi_md as MD

pytestmark = pytest.mark.gpu

LF = (8, 8, 8, 12)  # coarse (4, 4, 4, 6)


def _coarse(L):
    return tuple(v // 2 for v in L)


def _vol(L):
    return int(np.prod(L))


def link(ctx, L, host):
    return M.Field(ctx, L, LINK18, F64).upload(host)


def herm(orc, a):
    """(V, 4, 18) reals of a full 3x3 -> complex (V, 4, 3, 3)"""
    return orc.as_cplx(a)


def _random_state(ctx, orc, seed, L=LF):
    U = orc.haar_links(L, seed)
    Uc = orc.haar_links(_coarse(L), seed + 1)
    xi = np.random.default_rng(seed).uniform(0.5, 1.5, size=(_vol(_coarse(L)), 4))
    return U, Uc, xi


def _smooth_state(ctx, orc, seed, L=LF, eps=0.25):
    U = orc.unit_links(L)
    orc.evolve_gauge_field(L, U, orc.init_momentum(L, seed, 0), eps, 0)
    mp = orc.multi(0.1, 1.0, 0.5)
    Uc = orc.multi_gb(L, U, mp)
    orc.evolve_gauge_field(_coarse(L), Uc, orc.init_momentum(_coarse(L), seed + 1, 0), 0.3, 0)
    return U, Uc, np.ones((_vol(_coarse(L)), 4))


def _dev(ctx, L, U, Uc, xi):
    return link(ctx, L, U), link(ctx, _coarse(L), Uc), M.Field(ctx, _coarse(L), REAL1, F64).upload(xi)


def test_real1_field_roundtrip(ctx):
    x = np.random.default_rng(1).normal(size=(_vol(_coarse(LF)), 4))
    f = M.Field(ctx, _coarse(LF), REAL1, F64).upload(x)
    np.testing.assert_array_equal(f.download(), x)
    with pytest.raises(M.MagError):
        M.Field(ctx, _coarse(LF), REAL1, M.F32)


def test_gb_and_xi_energy(ctx, orc):
    U, Uc, xi = _random_state(ctx, orc, 11)
    mp = MD.params(0.15, 1.0)
    omp = orc.multi(0.15, 1.0, 0.5)
    f, c, x = _dev(ctx, LF, U, Uc, xi)
    gb = MD.get_Gb(M.Field(ctx, _coarse(LF), LINK18, F64), f, mp).download()
    assert np.abs(gb - orc.multi_gb(LF, U, omp)).max() < 1e-13
    s, ref = MD.get_xi_energy(f, c, x, mp), orc.xi_energy(LF, U, Uc, xi, omp)
    assert abs(s - ref) < 1e-12 * abs(ref)


def test_get_Q(ctx, orc):
    """get_Q (field-multi.h:32-39) at every coarse link against the oracle."""
    U, Uc, xi = _random_state(ctx, orc, 13)
    f = link(ctx, LF, U)
    Q = MD.get_Q(M.Field(ctx, _coarse(LF), ALG8, F64), f, MD.params(0.15, 1.0)).download()
    assert np.abs(Q - orc.multi_Q(LF, U, orc.multi(0.15, 1.0, 0.5))).max() < 1e-13


@pytest.mark.parametrize("gauge", [M.Gauge(M.WILSON, 0.0), M.Gauge(M.IWASAKI, -0.331)])
def test_fine_force(ctx, orc, gauge):
    U, Uc, xi = _random_state(ctx, orc, 21)
    arg = M.Arg_chmc(beta=5.7 if gauge.type == M.WILSON else 2.6, gauge=gauge)
    mp = MD.params(0.15, 1.0)
    f, c, x = _dev(ctx, LF, U, Uc, xi)
    F = M.Field(ctx, LF, ALG8, F64)
    MD.get_Fforce(F, f, c, arg, x, mp)
    ref = orc.multi_fforce(LF, U, Uc, xi, orc.action(gauge.type, arg.beta, gauge.c1), orc.multi(0.15, 1.0, 0.5))
    d = np.abs(F.download() - ref).max()
    assert d < 1e-12 * np.abs(ref).max(), d
    # the stout term is present on the links of classes 1-4
    Fg = orc.get_force(LF, U, orc.action(gauge.type, arg.beta, gauge.c1))
    assert np.abs(ref - Fg).max() > 1e-3


def test_coarse_force(ctx, orc):
    U, _, _ = _random_state(ctx, orc, 31)
    P = orc.init_momentum(LF, 32, 0)
    mp = MD.params(0.15, 1.0)
    f = link(ctx, LF, U)
    p = M.Field(ctx, LF, ALG8, F64).upload(P)
    cf = MD.get_Cforce(M.Field(ctx, _coarse(LF), ALG8, F64), f, p, mp).download()
    ref = orc.multi_cforce(LF, U, P, orc.multi(0.15, 1.0, 0.5))
    assert np.abs(cf - ref).max() < 1e-12 * np.abs(ref).max()


def test_coarse_force_at_zero_Q(ctx, orc):
    """Unit fine links give Q = 0 exactly; DexpiQ then returns zero, as the
    reference's does (field-multi.h:100-102), leaving e^{iQ} DU2 U2^+."""
    U = orc.unit_links(LF)
    P = orc.init_momentum(LF, 33, 0)
    mp = MD.params(0.1, 1.0)
    cf = MD.get_Cforce(M.Field(ctx, _coarse(LF), ALG8, F64), link(ctx, LF, U),
                       M.Field(ctx, LF, ALG8, F64).upload(P), mp).download()
    ref = orc.multi_cforce(LF, U, P, orc.multi(0.1, 1.0, 0.5))
    assert np.abs(cf - ref).max() < 1e-13


def test_heatbath_same_decisions(ctx, orc):
    U, Uc, _ = _random_state(ctx, orc, 41)
    mp = MD.params(0.1, 1.0)
    omp = orc.multi(0.1, 1.0, 0.5)
    f, c = link(ctx, LF, U), link(ctx, _coarse(LF), Uc)
    ref = Uc.copy()
    for t in range(3):
        MD.heatbath(c, f, mp, 42, t)
        orc.multi_heatbath(LF, ref, U, omp, 42, t)
    out = c.download()
    assert np.abs(out - ref).max() < 1e-13
    moved = np.any(np.abs(out - Uc) > 0, axis=2).mean()
    assert 0.5 < moved < 1.0


def test_zeta_T(ctx, orc):
    U, Uc, _ = _smooth_state(ctx, orc, 51)
    P = orc.init_momentum(LF, 52, 0)
    mp = MD.params(0.1, 1.0)
    f, c = link(ctx, LF, U), link(ctx, _coarse(LF), Uc)
    MD.force_gradient_integrator_zeta_T(f, M.Field(ctx, LF, ALG8, F64).upload(P), c, 0.2, M.Arg_chmc(), mp)
    orc.multi_zeta_T(LF, U, P, Uc, 0.2, 0, orc.multi(0.1, 1.0, 0.5))
    assert np.abs(f.download() - U).max() < 1e-12
    assert np.abs(c.download() - Uc).max() < 1e-12


@pytest.mark.parametrize("which", [MD.ZETA_H, MD.FG, MD.NESTED])
def test_integrators(ctx, orc, which):
    U, Uc, xi = _smooth_state(ctx, orc, 61)
    P = orc.init_momentum(LF, 62, 0)
    arg = M.Arg_chmc(beta=5.7, trajectory_length=2, dt=0.05)
    mp = MD.params(0.1, 1.0)
    f, c, x = _dev(ctx, LF, U, Uc, xi)
    p = M.Field(ctx, LF, ALG8, F64).upload(P)
    if which == MD.ZETA_H:
        dev = MD.force_gradient_integrator_zeta_H(f, p, c, x, arg, mp)
    elif which == MD.FG:
        dev = MD.force_gradient_integrator(f, p, arg, 2, c, x, mp)
    else:
        dev = MD.nested_integrator(f, p, arg, c, x, mp)
    odev = orc.multi_integrate(LF, U, P, Uc, xi, orc.action(0, 5.7), orc.multi(0.1, 1.0, 0.5), which, 2, 0.05)
    assert np.abs(f.download() - U).max() < 1e-11
    assert np.abs(c.download() - Uc).max() < 1e-11
    assert np.abs(p.download() - P).max() < 1e-10
    assert np.all(dev < 1e-12) and np.all(odev < 1e-12)


def test_hamiltonian(ctx, orc):
    U, Uc, xi = _random_state(ctx, orc, 71)
    P = orc.init_momentum(LF, 72, 0)
    for gauge, beta in [(M.Gauge(M.WILSON, 0.0), 5.7), (M.Gauge(M.IWASAKI, -0.331), 2.6)]:
        arg = M.Arg_chmc(beta=beta, gauge=gauge)
        mp = MD.params(0.15, 0.8)
        f, c, x = _dev(ctx, LF, U, Uc, xi)
        h, parts = MD.get_hamiltonian_multi(f, M.Field(ctx, LF, ALG8, F64).upload(P), arg, c, x, mp)
        oh, oparts = orc.multi_hamiltonian(LF, U, P, Uc, xi, orc.action(gauge.type, beta, gauge.c1),
                                           orc.multi(0.15, 0.8, 0.5))
        assert np.abs(parts - oparts).max() < 1e-12 * np.abs(oparts).max()
        assert abs(h - oh) < 1e-12 * abs(oh)


def test_run_hmc_multi_matches_oracle(ctx, orc):
    U, Uc, _ = _smooth_state(ctx, orc, 81)
    arg = M.Arg_chmc(beta=5.7, trajectory_length=3, dt=0.08, num_forced_accept_step=1, num_trajectory=4)
    mp = MD.params(0.1, 1.0)
    uf, uc, res = MD.run_hmc_multi(ctx, LF, U, arg, Uc, mp, seed=83)
    V, Vc = U.copy(), Uc.copy()
    h = orc.hmc(trajectory_length=3, dt=0.08, num_forced_accept=1, seed=83)
    nacc, ores = orc.run_hmc_multi(LF, V, Vc, orc.action(0, 5.7), h, orc.multi(0.1, 1.0, 0.5), 0, 4)
    assert [r.accept for r in res] == [r.accept for r in ores]
    for r, o in zip(res, ores):
        assert abs(r.dh - o.dh) < 1e-8, (r.dh, o.dh)
        assert abs(r.h_old - o.h_old) < 1e-12 * abs(o.h_old)
        assert abs(r.avg_plaq_fine - o.plaq_fine) < 1e-12
        assert abs(r.avg_plaq_coarse - o.plaq_coarse) < 1e-12
        assert r.status == 0
    assert np.abs(uf - V).max() < 1e-10
    assert np.abs(uc - Vc).max() < 1e-10


def test_checks(ctx):
    f = M.Field(ctx, LF, LINK18, F64).unit()
    c = M.Field(ctx, _coarse(LF), LINK18, F64).unit()
    x = MD.xi_field(ctx, LF, 1.0)
    mp = MD.params(0.1, 1.0)
    with pytest.raises(M.MagError):  # smeared action
        MD.get_Fforce(M.Field(ctx, LF, ALG8, F64), f, c, M.Arg_chmc(n_stout=1), x, mp)
    with pytest.raises(M.MagError):  # coarse extents
        MD.get_xi_energy(f, M.Field(ctx, LF, LINK18, F64).unit(), x, mp)
    with pytest.raises(M.MagError):  # xi of the wrong kind
        MD.get_xi_energy(f, c, c, mp)
    with pytest.raises(M.MagError):  # odd fine extents
        MD.get_Gb(c, M.Field(ctx, (8, 8, 8, 9), LINK18, F64).unit(), mp)
    # unit links: Gb = 1, S_xi = -(1 + alpha) 3 sum xi
    assert abs(MD.get_xi_energy(f, c, x, mp) + 1.5 * 3 * 4 * _vol(_coarse(LF))) < 1e-9


# --------------------------------------------------- properties at 16^4
LB = (16, 16, 16, 16)


def _dH(ctx, orc, n, tau, U0, P0, C0, xi):
    arg = M.Arg_chmc(beta=5.7, trajectory_length=n, dt=tau / n)
    mp = MD.params(0.1, 1.0)
    f, c, x = _dev(ctx, LB, U0, C0, xi)
    p = M.Field(ctx, LB, ALG8, F64).upload(P0)
    h0, _ = MD.get_hamiltonian_multi(f, p, arg, c, x, mp)
    MD.force_gradient_integrator_zeta_H(f, p, c, x, arg, mp)
    h1, _ = MD.get_hamiltonian_multi(f, p, arg, c, x, mp)
    return h1 - h0


def test_zeta_H_conserves_H_16(ctx, orc):
    U, Uc, xi = _smooth_state(ctx, orc, 91, L=LB)
    P = orc.init_momentum(LB, 92, 0)
    d = [abs(_dH(ctx, orc, n, 0.4, U, P, Uc, xi)) for n in (2, 4, 8)]
    assert d[0] / d[1] > 3.0 and d[1] / d[2] > 3.0, d


def test_heatbath_one_link_distribution_16(ctx, orc):
    """The heat bath's stationary one-link law: <Re Tr(C Gb^+)> under
    exp((1+alpha) XI0 Re Tr(C Gb^+)) dHaar, from the Weyl integral."""
    from test_multi_oracle import _weyl_mean_retr
    U = orc.haar_links(LB, 93)
    mp = MD.params(0.1, 1.2)
    f = link(ctx, LB, U)
    c = M.Field(ctx, _coarse(LB), LINK18, F64).haar(94)
    gb = MD.get_Gb(M.Field(ctx, _coarse(LB), LINK18, F64), f, mp).download()
    G = orc.as_cplx(gb)
    vals = []
    for t in range(400):
        MD.heatbath(c, f, mp, 95, t)
        if t >= 150 and t % 10 == 0:
            W = orc.as_cplx(c.download()) @ np.conj(np.swapaxes(G, -1, -2))
            vals.append(np.trace(W, axis1=-2, axis2=-1).real.mean())
    ref = _weyl_mean_retr(1.5 * 1.2)
    assert abs(np.mean(vals) - ref) < 0.02, (np.mean(vals), ref)
