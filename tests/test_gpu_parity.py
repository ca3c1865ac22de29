"""GPU parity: the gfx950 HIP library (through its C ABI) against the FP64
CPU oracle on the same seeded inputs.

Tolerances (FP64 path): per-matrix functions 1e-14 absolute; forces and
smeared fields 1e-12 relative (max-norm of the difference over max-norm of
the field); plaquette, rectangle, action and Hamiltonian 1e-13 relative;
identical accept/reject and demon sequences. FP32 molecular dynamics: the
errors against the FP64 oracle are asserted at the bounds stated per test.
"""
import numpy as np
import pytest

import mag_amd as M
from mag_amd import F32, F64, LINK12, LINK18, ALG8

pytestmark = pytest.mark.gpu

L4 = (4, 4, 4, 4)
L8 = (8, 8, 8, 8)


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def links(ctx, L, host, prec=F64):
    kind = LINK18 if prec == F64 else LINK12
    return M.Field(ctx, L, kind, prec).upload(host)


def alg(ctx, L, host=None, prec=F64):
    f = M.Field(ctx, L, ALG8, prec)
    if host is not None:
        f.upload(host)
    return f


def herm(orc, rng, scale, n):
    T = orc.generators()
    return np.einsum("na,aij->nij", rng.normal(size=(n, 8)) * scale, T)


# ------------------------------------------------------------------ per link
# FP32 kernels take exp(iQ) and Lambda coefficients from the Cayley-Hamilton
# reduced series for c1 = Tr Q^2 / 2 <= 1 and from the FP64 closed forms above
# that: the scales put the batches on either side ("mixed" alternates them
# lane by lane, so one wavefront takes both branches).
def scaled_herm(orc, rng, scale, n):
    if scale == "mixed":
        s = np.where(np.arange(n) % 2 == 0, 0.05, 1.5)
        return herm(orc, rng, 1.0, n) * s[:, None, None]
    return herm(orc, rng, scale, n)


@pytest.mark.parametrize("scale", [1e-5, 1e-2, 0.3, 1.5, "mixed"])
def test_expiQ_batch(ctx, orc, scale):
    rng = np.random.default_rng(1)
    Q = scaled_herm(orc, rng, scale, 4096)
    ref = orc.expiQ(Q)
    big = 1.5 if scale == "mixed" else scale
    assert np.abs(M.expiQ(ctx, Q, F64) - ref).max() < 1e-14 * max(1, big)
    assert np.abs(M.expiQ(ctx, Q, F32) - ref).max() < 2e-6


@pytest.mark.parametrize("scale", [1e-3, 0.05, 0.4, 1.5, "mixed"])
def test_compute_Lambda_batch(ctx, orc, scale):
    rng = np.random.default_rng(2)
    Q = scaled_herm(orc, rng, scale, 2048)
    S = rng.normal(size=(2048, 3, 3)) + 1j * rng.normal(size=(2048, 3, 3))
    U = orc.as_cplx(orc.haar_links(L8, 3)).reshape(-1, 3, 3)[:2048]
    ref = orc.compute_Lambda(Q, S, U)
    # The closed-form b_2j lose digits as c1 -> 0 (the oracle's too), so the
    # FP64 comparison is relative to the size of Lambda at each scale.
    assert rel(M.compute_Lambda(ctx, Q, S, U, F64), ref) < (1e-12 if scale != 1e-3 else 1e-9)
    tol32 = 5e-5 if scale in (1.5, "mixed") else 5e-6
    assert rel(M.compute_Lambda(ctx, Q, S, U, F32), ref) < tol32


@pytest.mark.parametrize("scale", [1e-3, 0.05, 0.4, 1.5, "mixed"])
def test_DexpiQ_batch(ctx, orc, scale):
    """DexpiQ (field-multi.h:90-178) through its batch entry point against the
    oracle (itself pinned by finite differences of expm); zero at Q = 0."""
    rng = np.random.default_rng(5)
    Q = scaled_herm(orc, rng, scale, 2048)
    DQ = herm(orc, rng, 1.0, 2048)
    ref = orc.DexpiQ(Q, DQ)
    assert rel(M.DexpiQ(ctx, Q, DQ, F64), ref) < (1e-12 if scale != 1e-3 else 1e-9)
    assert rel(M.DexpiQ(ctx, Q, DQ, F32), ref) < (5e-5 if scale in (1.5, "mixed") else 5e-6)
    assert not M.DexpiQ(ctx, np.zeros((4, 3, 3)), DQ[:4], F64).any()


def test_chair_and_twist_staples(ctx, orc):
    """chair/twist_staple_dagger (field-staple.h:141-165) for every link against
    the oracle's per-site sums (undaggered, defects kept)."""
    L = (4, 4, 4, 6)
    U = orc.haar_links(L, 17)
    dU = links(ctx, L, U)
    for dev, ref in ((M.chair_staple_dagger, orc.chair_staple_dagger),
                     (M.twist_staple_dagger, orc.twist_staple_dagger)):
        got = orc.as_cplx(dev(M.Field(ctx, L, LINK18, F64), dU).download())
        for site in (0, 5, 100, 383):
            for mu in range(4):
                assert np.abs(got[site, mu] - ref(L, U, site, mu)).max() < 1e-13


def test_upload_download_roundtrip(ctx, orc):
    U = orc.haar_links(L4, 4)
    assert np.array_equal(links(ctx, L4, U).download(), U)
    d32 = links(ctx, L4, U, F32).download()
    assert np.abs(d32 - U).max() < 5e-7  # 12-real FP32 storage, row 2 rebuilt
    P = orc.init_momentum(L4, 4, 0)
    assert np.abs(alg(ctx, L4, P).download() - P).max() < 1e-15


def test_haar_and_momentum_streams_match_oracle(ctx, orc):
    U = M.Field(ctx, L8, LINK18, F64).haar(1234)
    assert np.abs(U.download() - orc.haar_links(L8, 1234)).max() < 1e-13
    P = alg(ctx, L8)
    M.init_momentum(P, 1234, 7)
    assert np.abs(P.download() - orc.init_momentum(L8, 1234, 7)).max() < 1e-13


def test_reductions(ctx, orc):
    U = orc.haar_links(L8, 1234)
    d = links(ctx, L8, U)
    assert rel(M.total_plaq(d), orc.total_plaq(L8, U)) < 1e-13
    assert rel(M.total_rectangular(d), orc.total_rectangular(L8, U)) < 1e-13
    assert rel(M.avg_plaquette(d), orc.avg_plaquette(L8, U)) < 1e-13
    for r, T in [((1, 0, 0), 1), ((2, 1, 0), 2), ((1, 1, 1), 3)]:
        assert abs(M.avg_wilson_loop(d, r, T) - orc.avg_wilson_loop(L8, U, r, T)) < 1e-13
    unit = M.Field(ctx, L8, LINK18, F64).unit()
    assert M.total_plaq(unit) == 18 * 8 ** 4
    assert M.avg_wilson_loop(unit, (4, 0, 0), 4) == 3.0


ACTIONS = {
    "wilson": dict(type=0, beta=6.0),
    "iwasaki": dict(type=1, beta=2.6, c1=-0.331),
    "dbw2": dict(type=1, beta=1.04, c1=-1.4069),
    "iwasaki_mag2": dict(type=1, beta=2.6, c1=-0.331, mag=2),
    "wilson_stout1": dict(type=0, beta=5.8, n_stout=1, rho=0.1),
    "iwasaki_stout3": dict(type=1, beta=2.6, c1=-0.331, n_stout=3, rho=0.1),
    "wilson_stout2_mag2": dict(type=0, beta=6.0, mag=2, n_stout=2, rho=0.12),
}


@pytest.mark.parametrize("name", list(ACTIONS))
def test_force_fp64(ctx, orc, name):
    a = orc.action(**ACTIONS[name])
    L = L8 if a.n_stout == 0 else L4
    U = orc.haar_links(L, 17)
    ref = orc.get_force(L, U, a)
    F = alg(ctx, L)
    M.get_force(F, links(ctx, L, U), M._lib.Action(a.type, a.beta, a.c1, a.mag, a.n_stout, a.rho))
    assert rel(F.download(), ref) < 1e-12


@pytest.mark.parametrize("name", ["wilson", "iwasaki", "iwasaki_stout3"])
def test_force_fp32_error(ctx, orc, name):
    """FP32 molecular-dynamics force against the FP64 oracle on the same (FP32-rounded) links."""
    a = orc.action(**ACTIONS[name])
    L = L8 if a.n_stout == 0 else L4
    U32 = links(ctx, L, orc.haar_links(L, 19), F32)
    U = U32.download()  # the FP32 links, exactly, as FP64
    ref = orc.get_force(L, U, a)
    F = alg(ctx, L, prec=F32)
    M.get_force(F, U32, M._lib.Action(a.type, a.beta, a.c1, a.mag, a.n_stout, a.rho))
    err = rel(F.download(), ref)
    assert err < 2e-5, err


# Volumes of the serving stream and non-cubic shapes. The stencil tile order
# has three paths: XCD slabs (an x3 slice of whole 256-site blocks and L3
# divisible by 8: 16x16x4x8 with whole (x0, x1) planes, 8x8x4x16 with a
# 64-site plane, where FP64 takes the next path), XCD-interleaved tiles
# (block count a multiple of 8: 6^4, 4x6x8x10, 12^4, with partial last
# blocks in the first two) and plain order (4^4). All must give the oracle's
# force and Hamiltonian.
@pytest.mark.parametrize("L", [(4, 4, 4, 4), (6, 6, 6, 6), (4, 6, 8, 10), (12, 12, 12, 12), (16, 16, 4, 8),
                               (8, 8, 4, 16)])
def test_force_volume_sweep(ctx, orc, L):
    a = orc.action(**ACTIONS["iwasaki_stout3"])
    U = orc.haar_links(L, 23)
    ref = orc.get_force(L, U, a)
    act = M._lib.Action(a.type, a.beta, a.c1, a.mag, a.n_stout, a.rho)
    F = alg(ctx, L)
    M.get_force(F, links(ctx, L, U), act)
    assert rel(F.download(), ref) < 1e-12
    U32 = links(ctx, L, U, F32)
    ref32 = orc.get_force(L, U32.download(), a)
    F32f = alg(ctx, L, prec=F32)
    M.get_force(F32f, U32, act)
    assert rel(F32f.download(), ref32) < 2e-5
    P = orc.init_momentum(L, 24, 0)
    h, parts = M.get_hamiltonian(links(ctx, L, U), alg(ctx, L, P), act)
    h_ref, parts_ref = orc.hamiltonian(L, U, P, a)
    assert rel(h, h_ref) < 1e-13
    assert rel(parts, parts_ref) < 1e-13


@pytest.mark.parametrize("mag", [0, 2])
def test_evolve_gauge_and_momentum(ctx, orc, mag):
    U = orc.haar_links(L8, 23)
    P = orc.init_momentum(L8, 23, 0)
    dU, dP = links(ctx, L8, U), alg(ctx, L8, P)
    M.evolve_gauge_field(dU, dP, 0.137, M._lib.Action(0, 6.0, 0.0, mag, 0, 0.0))
    Ur = U.copy()
    orc.evolve_gauge_field(L8, Ur, P, 0.137, mag)
    assert np.abs(dU.download() - Ur).max() < 1e-13
    F = orc.get_force(L8, U, orc.action(1, 2.6, -0.331))
    dF = alg(ctx, L8, F)
    M.evolve_momentum(dP, dF, 0.25)
    Pr = P.copy()
    orc.evolve_momentum(L8, Pr, F, 0.25)
    assert np.abs(dP.download() - Pr).max() < 1e-13


def test_stout_smear_fp64(ctx, orc):
    U = orc.haar_links(L4, 29)
    out = M.Field(ctx, L4, LINK18, F64)
    M.stout_smear(out, links(ctx, L4, U), 0.1, 3)
    assert rel(out.download(), orc.stout_smear(L4, U, 0.1, 3)) < 1e-12


@pytest.mark.parametrize("name", ["wilson", "iwasaki_stout3", "iwasaki_mag2"])
def test_hamiltonian(ctx, orc, name):
    a = orc.action(**ACTIONS[name])
    L = L4 if a.n_stout else L8
    U, P = orc.haar_links(L, 31), orc.init_momentum(L, 31, 0)
    h_ref, parts_ref = orc.hamiltonian(L, U, P, a)
    h, parts = M.get_hamiltonian(links(ctx, L, U), alg(ctx, L, P),
                                 M._lib.Action(a.type, a.beta, a.c1, a.mag, a.n_stout, a.rho))
    assert rel(h, h_ref) < 1e-13
    assert rel(parts, parts_ref) < 1e-13


def test_fg_trajectory_c1(ctx, orc):
    """Config C1: 8^4 Wilson beta=6.0, Haar links seed 1234, FG tau=1 with 8 steps."""
    a = orc.action(0, 6.0)
    U, P = orc.haar_links(L8, 1234), orc.init_momentum(L8, 1234, 0)
    h0, _ = orc.hamiltonian(L8, U, P, a)
    Ur, Pr = U.copy(), P.copy()
    dev_ref = orc.force_gradient_integrator(L8, Ur, Pr, a, 8, 0.125, reunit=True)
    h1, _ = orc.hamiltonian(L8, Ur, Pr, a)
    dU, dP = links(ctx, L8, U), alg(ctx, L8, P)
    arg = M.Arg_chmc(trajectory_length=8, dt=0.125, beta=6.0)
    dev = M.force_gradient_integrator(dU, dP, arg)
    g0, _ = M.get_hamiltonian(dU, dP, arg)
    assert rel(dU.download(), Ur) < 1e-10
    assert rel(dP.download(), Pr) < 1e-10
    assert abs(g0 - h1) < 1e-9 * abs(h1)
    assert abs((g0 - h0) - (h1 - h0)) < 1e-6 * max(1.0, abs(h1 - h0))
    assert dev < 1e-12 and dev_ref < 1e-12


# FP64 links stored as 12 reals (row 2 rebuilt on load): the same parity bar
# as the 18-real FP64 path on exact SU(3) (Haar) links.
@pytest.mark.parametrize("name", ["wilson", "iwasaki", "iwasaki_stout3", "iwasaki_mag2"])
def test_force_fp64_12real(ctx, orc, name):
    a = orc.action(**ACTIONS[name])
    L = L8 if a.n_stout == 0 else L4
    U = orc.haar_links(L, 17)
    ref = orc.get_force(L, U, a)
    F = alg(ctx, L)
    U12 = M.Field(ctx, L, LINK12, F64).upload(U)
    assert rel(U12.download(), U) < 1e-13  # row 2 rebuilt as conj(row0 x row1)
    M.get_force(F, U12, M._lib.Action(a.type, a.beta, a.c1, a.mag, a.n_stout, a.rho))
    assert rel(F.download(), ref) < 1e-12


def test_fg_trajectory_stout_fp64_12real(ctx, orc):
    """Iwasaki + 3-stout FG trajectory with reunitarization on FP64 12-real links, and its Hamiltonian."""
    a = orc.action(1, 2.6, -0.331, n_stout=3, rho=0.1)
    U, P = orc.haar_links(L4, 38), orc.init_momentum(L4, 38, 0)
    Ur, Pr = U.copy(), P.copy()
    orc.force_gradient_integrator(L4, Ur, Pr, a, 4, 0.1, reunit=True)
    h_ref, _ = orc.hamiltonian(L4, Ur, Pr, a)
    dU, dP = M.Field(ctx, L4, LINK12, F64).upload(U), alg(ctx, L4, P)
    arg = M.Arg_chmc(trajectory_length=4, dt=0.1, beta=2.6, gauge=M.Gauge(1, -0.331), n_stout=3, rho=0.1)
    M.force_gradient_integrator(dU, dP, arg)
    assert rel(dU.download(), Ur) < 1e-10
    assert rel(dP.download(), Pr) < 1e-10
    h, _ = M.get_hamiltonian(dU, dP, arg)
    assert abs(h - h_ref) < 1e-9 * abs(h_ref)


def test_accept_sequence_fp64_12real(ctx, orc):
    """The 20-trajectory accept sequence of the 18-real FP64 chain, with 12-real FP64 links."""
    ref, res18, _, _ = _chain_vs_oracle(ctx, orc, L4, dict(type=1, beta=2.6, c1=-0.331, n_stout=1, rho=0.1),
                                        dict(trajectory_length=4, dt=0.25, num_forced_accept=3, seed=98), 20)
    arg = M.Arg_chmc(trajectory_length=4, dt=0.25, beta=2.6, gauge=M.Gauge(1, -0.331), n_stout=1, rho=0.1,
                     num_forced_accept_step=3)
    ch = M.Chain(ctx, L4, arg, seed=98, precision=F64, recon12=True)
    ch.set_links(orc.haar_links(L4, 1234))
    res = ch.run(20)
    assert [r.accept for r in res] == [r.accept for r in ref] == [r.accept for r in res18]
    assert 0 < sum(r.accept for r in ref[3:]) < 17
    for r, q in zip(res, ref):
        assert abs(r.h_old - q.h_old) < 1e-12 * abs(q.h_old)
        assert abs(r.dh - q.dh) < 1e-8 * max(1, abs(q.dh))


def test_fg_trajectory_stout_fp64(ctx, orc):
    a = orc.action(1, 2.6, -0.331, n_stout=2, rho=0.1)
    U, P = orc.haar_links(L4, 37), orc.init_momentum(L4, 37, 0)
    Ur, Pr = U.copy(), P.copy()
    orc.force_gradient_integrator(L4, Ur, Pr, a, 3, 0.1, reunit=False)
    dU, dP = links(ctx, L4, U), alg(ctx, L4, P)
    arg = M.Arg_chmc(trajectory_length=3, dt=0.1, beta=2.6, gauge=M.Gauge(1, -0.331), n_stout=2, rho=0.1)
    M.force_gradient_integrator(dU, dP, arg, reunit=False)
    assert rel(dU.download(), Ur) < 1e-10
    assert rel(dP.download(), Pr) < 1e-10


def _chain_vs_oracle(ctx, orc, L, act_kw, h_kw, ntraj, mode=0, emax=0.0, links_seed=1234, U0=None):
    a = orc.action(**act_kw)
    U = orc.haar_links(L, links_seed) if U0 is None else U0.copy()
    h = orc.hmc(**h_kw, accept_mode=mode, demon_emax=emax)
    ds = None
    if mode:
        # the chain's window is relative to the initial top-level action sums
        top = U if a.n_stout == 0 else orc.stout_smear(L, U, a.rho, a.n_stout)
        ds = np.array([-orc.total_plaq(L, top), -orc.total_rectangular(L, top), 0.0, 0.0])
    Ur = U.copy()
    nacc, ref, _ = orc.run_hmc(L, Ur, a, h, 0, ntraj, ds)
    arg = M.Arg_chmc(trajectory_length=h.trajectory_length, dt=h.dt, beta=a.beta, gauge=M.Gauge(a.type, a.c1),
                     mag=a.mag, n_stout=a.n_stout, rho=a.rho, num_forced_accept_step=h.num_forced_accept)
    ch = M.Chain(ctx, L, arg, seed=h.seed, precision=F64, accept_mode=mode, demon_emax=emax)
    ch.set_links(U)
    res = ch.run(ntraj)
    return ref, res, Ur, ch


def test_accept_sequence_20_trajectories(ctx, orc):
    ref, res, Ur, ch = _chain_vs_oracle(ctx, orc, L4, dict(type=0, beta=5.7),
                                        dict(trajectory_length=3, dt=1 / 3, num_forced_accept=3, seed=99), 20)
    assert [r.accept for r in res] == [r.accept for r in ref]
    assert 0 < sum(r.accept for r in ref[3:]) < 17  # both outcomes occur
    for r, q in zip(res, ref):
        assert r.die_roll == q.die_roll
        assert abs(r.h_old - q.h_old) < 1e-12 * abs(q.h_old)
        assert abs(r.dh - q.dh) < 1e-8 * max(1, abs(q.dh))
        assert abs(r.avg_plaq - q.avg_plaq) < 1e-12
    assert rel(ch.get_links(), Ur) < 1e-9


def test_demon_window_sequence(ctx, orc):
    """Constrained Metropolis + demon window (accept mode 2, field-demon.h:152-176)
    on an Iwasaki + 3-stout chain over 20 fixed-seed trajectories, from links
    thermalised by 30 forced-accept oracle trajectories (a hot start relaxes by
    far more than any window per trajectory). The window (+-30 in raw Re Tr
    sums) rejects trajectories Metropolis alone would accept, so the accept
    sequence and the demon energies exercise both rules; both must be
    identical to the oracle's."""
    a = orc.action(type=1, beta=2.6, c1=-0.331, n_stout=3, rho=0.1)
    U0 = orc.haar_links(L4, 1234)
    orc.run_hmc(L4, U0, a, orc.hmc(trajectory_length=6, dt=1 / 6, num_forced_accept=30, seed=11), 0, 30)
    ref, res, Ur, ch = _chain_vs_oracle(ctx, orc, L4, dict(type=1, beta=2.6, c1=-0.331, n_stout=3, rho=0.1),
                                        dict(trajectory_length=6, dt=1 / 6, num_forced_accept=0, seed=3), 20,
                                        mode=2, emax=30.0, U0=U0)
    acc = [r.accept for r in ref]
    assert [r.accept for r in res] == acc
    metro = [q.die_roll < np.exp(-q.dh) for q in ref]
    assert 0 < sum(acc) < 20 and any(m and not x for m, x in zip(metro, acc))  # the window rejects
    assert len({(round(q.demon_e[0], 9), round(q.demon_e[1], 9)) for q in ref}) > 2  # the demons move
    for r, q in zip(res, ref):
        assert abs(r.demon_e[0] - q.demon_e[0]) < 1e-9 and abs(r.demon_e[1] - q.demon_e[1]) < 1e-9
        assert abs(r.h_old - q.h_old) < 1e-12 * abs(q.h_old)
        assert abs(r.dh - q.dh) < 1e-8 * max(1, abs(q.dh))
    assert rel(ch.get_links(), Ur) < 1e-9


def test_reversibility_fp64(ctx, orc):
    arg = M.Arg_chmc(trajectory_length=8, dt=0.125, beta=2.6, gauge=M.Gauge(1, -0.331), n_stout=1, rho=0.1)
    ch = M.Chain(ctx, L4, arg, seed=5)
    ch.set_links(orc.haar_links(L4, 5))
    dev, dh = ch.reversibility()
    assert dev < 1e-10
    assert abs(dh) < 1e-8


def test_fp32_chain_errors(ctx, orc):
    """FP32-MD / FP64-dH chain against the FP64 chain from the same start and seed."""
    arg = M.Arg_chmc(trajectory_length=8, dt=0.125, beta=2.6, gauge=M.Gauge(1, -0.331), n_stout=1, rho=0.1,
                     num_forced_accept_step=2)
    U = orc.haar_links(L8, 43)
    c64 = M.Chain(ctx, L8, arg, seed=43, precision=F64)
    c32 = M.Chain(ctx, L8, arg, seed=43, precision=F32)
    c64.set_links(U)
    c32.set_links(U)
    r64, r32 = c64.run(3), c32.run(3)
    for a, b in zip(r64, r32):
        assert abs(a.h_old - b.h_old) < 1e-5 * abs(a.h_old)
        assert abs(a.dh - b.dh) < 5e-2 * max(1.0, abs(a.dh))


def test_demon_microcanonical_sequence(ctx, orc):
    U = orc.haar_links(L4, 111)
    Ur = U.copy()
    E_ref, ar_ref = orc.demon_microcanonical(L4, Ur, 1, 40.0, 5)
    d = links(ctx, L4, U)
    E, ar = M.demon_microcanonical(d, 1, 40.0, 5)
    assert ar == ar_ref
    assert np.abs(E - E_ref).max() < 1e-9
    assert np.abs(d.download() - Ur).max() < 1e-12


def test_bitwise_determinism(ctx, orc):
    arg = M.Arg_chmc(trajectory_length=4, dt=0.25, beta=2.6, gauge=M.Gauge(1, -0.331), n_stout=2, rho=0.1)
    U = orc.haar_links(L4, 47)
    outs = []
    for _ in range(2):
        ch = M.Chain(ctx, L4, arg, seed=47, precision=F32)
        ch.set_links(U)
        r = ch.run(2)
        outs.append((ch.get_links(), [x.h_new for x in r]))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]


def test_cpp_driver_runs():
    """The reference-style C++ driver (run_chmc call sequence through include/mag_amd.hpp)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "chmc_driver")
    assert os.path.exists(exe), "build with make -C mag_amd"
    out = subprocess.run([exe, "3"], capture_output=True, text=True, timeout=100)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = [l for l in out.stdout.splitlines() if l[:1].isdigit()]
    assert len(rows) == 3 and "Accept Rate" in out.stdout and "demon:" in out.stdout


def test_gpu_against_golden_fixtures(ctx, orc):
    """The GPU path against the committed fixtures (tests/golden/): C1 trajectory and 20-trajectory sequence."""
    import os
    G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    d = np.load(os.path.join(G, "c1_wilson_8x4.npz"))
    dU = M.Field(ctx, L8, LINK18, F64).haar(1234)
    assert abs(M.total_plaq(dU) - d["total_plaq"]) < 1e-12 * abs(d["total_plaq"])
    dP = alg(ctx, L8)
    M.init_momentum(dP, 1234, 0)
    arg = M.Arg_chmc(trajectory_length=8, dt=0.125, beta=6.0)
    h0, _ = M.get_hamiltonian(dU, dP, arg)
    assert abs(h0 - d["h_old"]) < 1e-12 * abs(d["h_old"])
    M.force_gradient_integrator(dU, dP, arg)
    h1, _ = M.get_hamiltonian(dU, dP, arg)
    assert abs(h1 - d["h_new"]) < 1e-9 * abs(d["h_new"])
    s = np.load(os.path.join(G, "accept_seq_wilson_4x4.npz"))
    arg = M.Arg_chmc(trajectory_length=3, dt=1 / 3, beta=5.7, num_forced_accept_step=3)
    ch = M.Chain(ctx, L4, arg, seed=99)
    ch.haar(1234)
    res = ch.run(20)
    assert [r.accept for r in res] == list(s["accept"])
    assert np.allclose([r.dh for r in res], s["dh"], rtol=1e-6, atol=1e-9)


def test_c2_fp32_md_against_fp64_16x4(ctx, orc):
    """16^4 Iwasaki (c1 = -0.331, beta = 2.6) + 3-stout (rho = 0.1): Haar seed 7,
    20 FP64 thermalisation trajectories, then 50 trajectories (tau = 1, FG,
    10 steps) twice from the same start and seed, all-FP64 and FP32-MD /
    FP64-dH. Force and plaquette of the thermalised links against the FP64
    oracle; paired Metropolis probabilities matched within 2 %; dH agreement;
    <exp(-dH)> = 1."""
    L16 = (16, 16, 16, 16)
    kw = dict(trajectory_length=10, dt=0.1, beta=2.6, gauge=M.Gauge(M.IWASAKI, -0.331), n_stout=3, rho=0.1)
    therm = M.Chain(ctx, L16, M.Arg_chmc(num_forced_accept_step=20, **kw), seed=7, precision=F64)
    therm.haar(7)
    therm.run(20)
    U = therm.get_links()
    a = orc.action(type=1, beta=2.6, c1=-0.331, n_stout=3, rho=0.1)
    act = M._lib.Action(1, 2.6, -0.331, 0, 3, 0.1)
    ref = orc.get_force(L16, U, a)
    F = alg(ctx, L16)
    M.get_force(F, links(ctx, L16, U), act)
    assert rel(F.download(), ref) < 1e-12
    U32 = links(ctx, L16, U, F32)
    F32f = alg(ctx, L16, prec=F32)
    M.get_force(F32f, U32, act)
    assert rel(F32f.download(), orc.get_force(L16, U32.download(), a)) < 2e-5
    assert rel(M.avg_plaquette(U32), orc.avg_plaquette(L16, U)) < 1e-7

    arg = M.Arg_chmc(**kw)
    res = {}
    for prec in (F64, F32):
        ch = M.Chain(ctx, L16, arg, seed=8, precision=prec)
        ch.set_links(U)
        res[prec] = ch.run(50)
    a64 = np.array([r.accept for r in res[F64]])
    a32 = np.array([r.accept for r in res[F32]])
    dh64 = np.array([r.dh for r in res[F64]])
    dh32 = np.array([r.dh for r in res[F32]])
    info = f"accept64={a64.astype(int).tolist()} accept32={a32.astype(int).tolist()} dh64={dh64.round(4).tolist()} " \
           f"dh32={dh32.round(4).tolist()}"
    assert a64.mean() > 0.2, info
    # The first trajectories start from the same links (to FP32 rounding) and
    # integrate the same momenta: there the FP32-MD dH and the Metropolis
    # probability min(1, exp(-dH)) must match the FP64 ones. Later the two
    # chains drift apart (molecular dynamics amplifies the rounding
    # differences from one trajectory to the next) and become independent
    # samples whose acceptance rates agree statistically.
    n = int(np.argmax(a64 != a32)) if np.any(a64 != a32) else len(a64)
    assert n >= 10, info
    assert np.abs(dh32[:10] - dh64[:10]).max() < 0.02, info
    # Matched acceptance, paired: over every trajectory of the two chains
    # that starts from the same state (all while the decisions agree, and the
    # first that differs), the Metropolis probabilities agree on average
    # within 2 % and one by one within 2 % over the first ten.
    n = min(n + 1, len(a64))
    p64, p32 = np.minimum(1, np.exp(-dh64[:n])), np.minimum(1, np.exp(-dh32[:n]))
    assert abs((p32 - p64).mean()) < 0.02, info
    assert np.abs(p32[:10] - p64[:10]).max() < 0.02, info
    w = np.exp(-dh64)
    assert abs(w.mean() - 1.0) < 4 * w.std(ddof=1) / np.sqrt(len(w)) + 1e-3, info


def test_c5_shape_64x4_dbw2_3stout_fp32(ctx):
    """64^4 DBW2 (c1 = -1.4069, beta = 1.04) + 3-stout FP32 chain from a Haar
    start (seed 99): one forced-accept force-gradient trajectory with 16 steps
    at the full volume (16.7 M sites, ~78 GB resident), then the 4x4 Wilson
    loop. Checks that every index and byte offset of the large lattice stays
    in range: finite energies, unitary links, and a Wilson loop that matches
    the FP64 re-evaluation of the same (FP32) links."""
    L64 = (64, 64, 64, 64)
    arg = M.Arg_chmc(trajectory_length=16, dt=1 / 16, beta=1.04, gauge=M.Gauge(M.IWASAKI, -1.4069), n_stout=3,
                     rho=0.1, num_forced_accept_step=1)
    ch = M.Chain(ctx, L64, arg, seed=99, precision=F32)
    ch.haar(99)
    r = ch.run(1)[0]
    assert np.isfinite(r.h_old) and np.isfinite(r.h_new) and r.accept == 1
    assert abs(r.dh) < 1e-3 * abs(r.h_old)
    assert r.reunit_dev < 1e-4
    U = ch.links()
    w32 = M.avg_wilson_loop(U, (4, 0, 0), 4)
    U64 = M.Field(ctx, L64, LINK18, F64).copy_from(U)
    w64 = M.avg_wilson_loop(U64, (4, 0, 0), 4)
    assert np.isfinite(w32) and abs(w32 - w64) < 1e-12
    assert abs(M.avg_plaquette(U64) - r.avg_plaq) < 1e-12


def test_async_reductions_match_synchronous(ctx, orc):
    """The _async reductions leave their results in device memory, in stream
    order. Read back, the sums equal the synchronous calls bit for bit; the
    potential and H, which the device combines from the same sums with
    fused multiply-adds, agree to rounding. (Device memory for the results
    comes straight from the HIP runtime.)"""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(64)) == 0
    try:
        U = links(ctx, L8, orc.haar_links(L8, 3))
        P = alg(ctx, L8)
        M.init_momentum(P, 5)
        arg = M.Arg_chmc(beta=2.6, gauge=M.Gauge(M.IWASAKI, -0.331))
        base = ptr.value
        M.total_plaq_async(U, base)
        M.total_rectangular_async(U, base + 8)
        M.get_hamiltonian_async(U, P, arg, base + 16)
        M.avg_wilson_loop_async(U, (2, 1, 0), 2, base + 40)
        ctx.synchronize()
        h = np.zeros(8)
        assert hip.hipMemcpy(h.ctypes.data_as(ctypes.c_void_p), ptr, ctypes.c_size_t(48), 2) == 0  # D2H
        H, parts = M.get_hamiltonian(U, P, arg)
        assert h[0] == M.total_plaq(U) and h[1] == M.total_rectangular(U)
        assert h[2] == parts[0]
        assert abs(h[3] - parts[1]) <= 1e-14 * abs(parts[1]) and abs(h[4] - H) <= 1e-14 * abs(H)
        assert h[5] == M.avg_wilson_loop(U, (2, 1, 0), 2)
    finally:
        hip.hipFree(ptr)


@pytest.mark.parametrize("prec", [F64, F32])
def test_fg_dh_scales_as_dt4(ctx, orc, prec):
    """Force-gradient integration on the device: from a smooth configuration
    (links e^{iQ} with |Q| ~ 0.3) and fixed momenta, halving the step at
    tau = 1 divides dH by 2^4 (Iwasaki + 1-stout). FP32 MD holds the ratio
    while dH stays well above its rounding noise."""
    rng = np.random.default_rng(4)
    T = orc.generators()
    q = rng.normal(size=(np.prod(L8) * 4, 8)) * 0.3
    U0 = orc.expiQ(np.einsum("na,aij->nij", q, T))
    U0 = orc.as_real(U0).reshape(-1, 4, 18)
    P0 = orc.init_momentum(L8, 77, 0)
    dh = []
    for n in (8, 16):
        arg = M.Arg_chmc(trajectory_length=n, dt=1.0 / n, beta=5.0, gauge=M.Gauge(M.IWASAKI, -0.331), n_stout=1,
                         rho=0.1)
        U, P = links(ctx, L8, U0, prec), alg(ctx, L8, P0, prec)
        h0, _ = M.get_hamiltonian(U, P, arg)
        M.force_gradient_integrator(U, P, arg, reunit=False)
        h1, _ = M.get_hamiltonian(U, P, arg)
        dh.append(h1 - h0)
    ratio = dh[0] / dh[1]
    assert 13.0 < ratio < 19.0, dh
