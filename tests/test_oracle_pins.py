"""Independent pins of the oracle functions that the finite-difference and
invariance tests cannot fix on their own:

- the stout forward map (isotropic Q and U' = e^{iQ} U for 1 and 3 levels)
  and the 2-link get_Q of the md:: path, against numpy restatements of
  hep-lat/0311018 eqs. 1-3 (np.roll staples, the paper's Q written out,
  scipy.linalg.expm) -- a force test differentiates whatever forward map the
  oracle uses, so a sign of rho or a factor inside the projection would pass
  it;
- avg_wilson_loop for loops with more than one nonzero spatial extent,
  against a brute-force numpy product along the path of field-wloop.h:26-56
  (the return leg in reverse spatial order), which a reversed return leg
  fails;
- the chair and twist path tables (field-staple.h:27-135) by their endpoint
  structure as the reference builds them (the backtracking chair bases of
  field-staple.h:39 and :63 close, the nu,-la,-la,nu and -nu,la,la,-nu bases
  of :47 and :55 end at x+mu+-2(nu-la), 72 paths per mu), and
  chair/twist_staple_dagger as the undaggered sum over them (:141-165).
"""
import itertools

import numpy as np
import pytest
import scipy.linalg as sl

L4 = (4, 4, 4, 4)


def lat(U, L):
    """(V,4,18) -> complex (L3, L2, L1, L0, 4, 3, 3); the axis of direction mu is 3 - mu."""
    return U.view(np.complex128).reshape(L[3], L[2], L[1], L[0], 4, 3, 3)


def unlat(X):
    return np.ascontiguousarray(X).reshape(-1, 4, 9).view(np.float64).reshape(-1, 4, 18)


def sh(A, mu, n=1):
    """A(x + n mu-hat) on the periodic lattice."""
    return np.roll(A, -n, axis=3 - mu)


def dg(A):
    return np.conj(np.swapaxes(A, -1, -2))


def tr(A):
    return np.trace(A, axis1=-2, axis2=-1)


def mp_Q(Omega):
    """hep-lat/0311018 eq. 2: Q = (i/2)(Omega^+ - Omega) - (i/6) Tr(Omega^+ - Omega)."""
    D = dg(Omega) - Omega
    return 0.5j * D - (1j / 6.0) * tr(D)[..., None, None] * np.eye(3)


def np_stout_level(X, rho):
    """One isotropic stout step, hep-lat/0311018 eqs. 1-3: C_mu = rho sum_nu of
    the upper and lower plaquette staples, Omega = C U^+, U' = exp(iQ) U."""
    Q = np.empty_like(X)
    for mu in range(4):
        Um = X[..., mu, :, :]
        C = np.zeros_like(Um)
        for nu in range(4):
            if nu == mu:
                continue
            Un = X[..., nu, :, :]
            C += Un @ sh(Um, nu) @ dg(sh(Un, mu))
            C += sh(dg(Un) @ Um @ sh(Un, mu), nu, -1)
        Q[..., mu, :, :] = mp_Q(rho * C @ dg(Um))
    return sl.expm(1j * Q) @ X, Q


def rand_links(orc, L, seed):
    return orc.haar_links(L, seed)


# ---------------------------------------------------------------- stout

@pytest.mark.parametrize("L", [(4, 4, 4, 4), (4, 4, 6, 4)])
def test_stout_Q_and_one_level_vs_numpy(orc, L):
    U = rand_links(orc, L, 21)
    X = lat(U, L)
    rho = 0.1
    Xn, Qn = np_stout_level(X, rho)
    Qo = lat(orc.stout_Q(L, U, rho), L)
    assert np.abs(Qo - Qn).max() < 1e-13
    Uo = lat(orc.stout_smear(L, U, rho, 1), L)
    assert np.abs(Uo - Xn).max() < 1e-13
    # a flipped sign of rho, or the projection without its factor 1/2, is far away
    Xm, _ = np_stout_level(X, -rho)
    assert np.abs(Uo - Xm).max() > 1e-2
    Qh = lat(orc.stout_Q(L, U, 2 * rho), L)
    assert np.abs(Qh - Qn).max() > 1e-2


def test_stout_three_levels_vs_numpy(orc):
    L = L4
    U = rand_links(orc, L, 22)
    X = lat(U, L)
    for _ in range(3):
        X, _ = np_stout_level(X, 0.1)
    Uo = lat(orc.stout_smear(L, U, 0.1, 3), L)
    assert np.abs(Uo - X).max() < 1e-13


def test_one_stout_level_raises_the_plaquette_of_a_haar_field(orc):
    L = (6, 6, 6, 6)
    U = rand_links(orc, L, 23)
    p0 = orc.avg_plaquette(L, U)
    p1 = orc.avg_plaquette(L, orc.stout_smear(L, U, 0.1, 1))
    pm = orc.avg_plaquette(L, unlat(np_stout_level(lat(U, L), -0.1)[0]))
    # Haar links: <plaq> ~ 0 with spread ~ 1/sqrt(6 V); smoothing lifts it well above that
    assert abs(p0) < 0.02 and p1 > p0 + 0.05 and pm < p1 - 0.05


def test_md_get_Q_and_Gb_vs_numpy(orc):
    """get_Q of field-multi.h:33-39 (2-link U2 = U_mu(x) U_mu(x+mu), six 2x1
    staples, Q = -i rho TA(stp U2^+)) and Gb = e^{iQ} U2 on the even sites,
    against numpy with the paper's Q."""
    L = (4, 4, 4, 8)
    U = rand_links(orc, L, 24)
    X = lat(U, L)
    rho = 0.1
    mp = orc.multi(rho, 1.0, 0.5)
    Gb_o = orc.multi_gb(L, U, mp)
    Lc = tuple(v // 2 for v in L)
    Gb_o = lat(Gb_o, Lc)
    Gb_n = np.empty_like(Gb_o)
    for mu in range(4):
        Um = X[..., mu, :, :]
        U2 = Um @ sh(Um, mu)
        stp = np.zeros_like(Um)
        for nu in range(4):
            if nu == mu:
                continue
            Un = X[..., nu, :, :]
            stp += Un @ sh(Um, nu) @ sh(sh(Um, nu), mu) @ dg(sh(Un, mu, 2))
            stp += sh(dg(Un) @ Um @ sh(Um, mu) @ sh(Un, mu, 2), nu, -1)
        Q = mp_Q(rho * stp @ dg(U2))
        G = sl.expm(1j * Q) @ U2
        Gb_n[..., mu, :, :] = G[::2, ::2, ::2, ::2]
    assert np.abs(Gb_o - Gb_n).max() < 1e-13


# ---------------------------------------------------------- Wilson loops

def np_wilson_loop(U, L, r, T, reverse_return=False):
    """field-wloop.h:26-98 by brute force: forward r[0] steps in direction 0,
    r[1] in 1, r[2] in 2, T in 3; back r[2] in 2, r[1] in 1, r[0] in 0 (or in
    the order 0, 1, 2 when reverse_return), T in 3; Re Tr averaged over the
    distinct permutations of sorted r and over sites."""
    X = lat(U, L)
    perms = sorted(set(itertools.permutations(sorted(r))))
    tot = 0.0
    for rr in perms:
        order = range(3) if reverse_return else range(2, -1, -1)
        steps = [(d, +1) for d in range(3) for _ in range(rr[d])] + [(3, +1)] * T
        steps += [(d, -1) for d in order for _ in range(rr[d])] + [(3, -1)] * T
        pos = np.zeros(4, dtype=int)
        M = np.broadcast_to(np.eye(3, dtype=complex), X.shape[:4] + (3, 3))
        for d, s in steps:
            if s > 0:
                A = X[..., d, :, :]
                for k in range(4):
                    A = sh(A, k, pos[k])
                M = M @ A
                pos[d] += 1
            else:
                pos[d] -= 1
                A = X[..., d, :, :]
                for k in range(4):
                    A = sh(A, k, pos[k])
                M = M @ dg(A)
        assert not pos.any()
        tot += tr(M).real.sum()
    return tot / (len(perms) * np.prod(L))


@pytest.mark.parametrize("L,r,T", [((4, 4, 4, 6), (2, 1, 0), 2), ((6, 6, 6, 6), (4, 0, 0), 4),
                                   ((5, 4, 6, 4), (1, 2, 1), 1), ((4, 4, 4, 4), (1, 1, 1), 2)])
def test_wilson_loop_vs_brute_force(orc, L, r, T):
    U = rand_links(orc, L, 31)
    # a smooth field (three stout levels) keeps large loops well away from zero
    U = orc.stout_smear(L, U, 0.12, 3)
    w = orc.avg_wilson_loop(L, U, r, T)
    ref = np_wilson_loop(U, L, r, T)
    assert abs(w - ref) < 1e-13 * max(1.0, abs(ref))
    if sum(v > 0 for v in r) > 1:
        # the return leg in the other spatial order is a different loop
        assert abs(np_wilson_loop(U, L, r, T, reverse_return=True) - ref) > 1e-6


# ------------------------------------------------------ chair / twist tables

def displacement(path):
    d = np.zeros(4, dtype=int)
    for c in path:
        if c < 4:
            d[c] += 1
        else:
            d[c - 4] -= 1
    return tuple(d)


def unit(mu):
    e = np.zeros(4, dtype=int)
    e[mu] = 1
    return e


@pytest.mark.parametrize("mu", range(4))
def test_chair_and_twist_table_structure(orc, mu):
    for twist in (False, True):
        paths = orc.staple_index(mu, twist)
        assert len(paths) == 72 and all(len(p) == 5 for p in paths)
        ends = {}
        for p in paths:
            # one forward mu step, never first or last; the rest a 4-step base in two transverse directions
            assert p.count(mu) == 1 and mu + 4 not in p and p[0] != mu and p[-1] != mu
            base = [c for c in p if c != mu]
            assert len({c % 4 for c in base}) == 2 and mu not in {c % 4 for c in base}
            d = displacement(p)
            ends[d] = ends.get(d, 0) + 1
            if not twist and d == tuple(unit(mu)):
                # the closed chair bases backtrack: a step followed by its inverse
                assert any(abs(base[i] - base[i + 1]) == 4 for i in range(3))
        xm = tuple(unit(mu))
        if twist:
            assert ends == {xm: 72}
        else:
            # 36 closed paths; for each ordered transverse pair (a, b) 6 paths end at x + mu + 2(e_a - e_b)
            expect = {xm: 36}
            for a, b in itertools.permutations([k for k in range(4) if k != mu], 2):
                expect[tuple(unit(mu) + 2 * unit(a) - 2 * unit(b))] = 6
            assert ends == expect


@pytest.mark.parametrize("mu", [0, 3])
def test_chair_and_twist_sums_are_the_table_products(orc, mu):
    """chair/twist_staple_dagger = the undaggered sum of the table's path
    products (field-staple.h:141-165), at two sites."""
    L = (4, 4, 4, 6)
    U = rand_links(orc, L, 41)
    for site in (0, 123):
        for twist in (False, True):
            s = sum(orc.path_product(L, U, list(np.unravel_index(site, L[::-1])[::-1]), p)
                    for p in orc.staple_index(mu, twist))
            f = orc.twist_staple_dagger if twist else orc.chair_staple_dagger
            got = f(L, U, site, mu)
            assert np.abs(got - s).max() < 1e-13
            assert np.abs(got - dg(s)).max() > 1e-3
