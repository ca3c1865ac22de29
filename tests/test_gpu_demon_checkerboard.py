"""Checkerboard demon (demon_microcanonical with a demon per link) on the
device: the oracle's class order gives the same links, demons and accept
counts (the classes commute, so thread order does not matter); at 16^4 the
closed-loop energies S_k + sum E_k are conserved and the windows hold; with
an unbounded window it ends where the sequential device demon ends."""
import numpy as np
import pytest

import mag_amd as M
from mag_amd import F64, LINK18, REAL1

pytestmark = pytest.mark.gpu


def _demons(ctx, L, E=None):
    vol = int(np.prod(L))
    E = np.zeros((vol, 4, 4)) if E is None else E
    return [M.Field(ctx, L, REAL1, F64).upload(np.ascontiguousarray(E[:, :, k])) for k in range(4)]


def _totals(demons):
    return np.array([d.download().sum() for d in demons])


# 8^4: 32 links per class, whole wavefronts of 8 links; 4x8x4x12: 12 links
# per class, the second wavefront half filled (its surplus lanes repeat the
# last link and must write nothing)
@pytest.mark.parametrize("L", [(8, 8, 8, 8), (4, 8, 4, 12)])
def test_matches_oracle(ctx, orc, L):
    U0 = orc.haar_links(L, 151)
    f = M.Field(ctx, L, LINK18, F64).upload(U0)
    d, (acc, rej) = M.demon_checkerboard(f, 2, 1.5, 152, _demons(ctx, L))
    V = U0.copy()
    E, (oacc, orej) = orc.demon_checkerboard(L, V, 2, 1.5, 152)
    assert (acc, rej) == (oacc, orej)
    assert np.abs(f.download() - V).max() < 1e-12
    for k in range(4):
        assert np.abs(d[k].download() - E[:, :, k]).max() < 1e-11


def test_conservation_16(ctx):
    L = (16, 16, 16, 16)
    f = M.Field(ctx, L, LINK18, F64).haar(161)
    s0 = np.array([-M.total_plaq(f), -M.total_rectangular(f)])
    d, (acc, rej) = M.demon_checkerboard(f, 3, 2.0, 162)
    s1 = np.array([-M.total_plaq(f), -M.total_rectangular(f)])
    tot = _totals(d)
    assert acc > 0 and rej > 0
    assert np.abs(s1 + tot[:2] - s0).max() < 1e-7
    assert max(np.abs(x.download()).max() for x in d) < 2.0


def test_unbounded_window_equals_sequential(ctx, orc):
    L = (4, 4, 4, 4)
    U0 = orc.haar_links(L, 171)
    a = M.Field(ctx, L, LINK18, F64).upload(U0)
    b = M.Field(ctx, L, LINK18, F64).upload(U0)
    d, (acc, rej) = M.demon_checkerboard(a, 1, 1e300, 172)
    Es, (acc_s, rej_s) = M.demon_microcanonical(b, 1, 1e300, 172)
    assert rej == 0 and rej_s == 0 and acc == acc_s
    assert np.abs(a.download() - b.download()).max() < 1e-13
    assert np.abs(_totals(d)[:2] - np.asarray(Es)[:2]).max() < 1e-9


def test_checks(ctx):
    f = M.Field(ctx, (8, 8, 8, 6), LINK18, F64).unit()
    with pytest.raises(M.MagError):
        M.demon_checkerboard(f, 1, 1.0, 1)
