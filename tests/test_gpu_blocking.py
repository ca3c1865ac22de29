"""Mag-blocked lattices on the device (coarsen, update_constrain,
check_constrained_plaquette, lattice_expansion; double_multigrid) against
the FP64 oracle and the identities that hold between them."""
import numpy as np
import pytest

import mag_amd as M
from mag_amd import F32, F64, LINK12, LINK18
from mag_amd.multigrid import double_multigrid

pytestmark = pytest.mark.gpu

LF, LC = (8, 8, 8, 8), (4, 4, 4, 4)


def field(ctx, L, host, prec=F64):
    return M.Field(ctx, L, LINK18 if prec == F64 else LINK12, prec).upload(host)


def test_primitives_match_oracle_fp64(ctx, orc):
    U, Uc = orc.haar_links(LF, 3), orc.haar_links(LC, 9)
    f, c = field(ctx, LF, U), M.Field(ctx, LC, LINK18, F64)
    M.coarsen_links(f, 2, c)
    assert np.abs(c.download() - orc.coarsen(LF, U, 2)).max() < 1e-14
    assert abs(M.check_constrained_plaquette(f, 2) - orc.constrained_plaquette(LF, U, 2)) < 1e-14
    c.upload(Uc)
    M.update_constrain(f, c, 2)
    V = U.copy()
    orc.update_constrain(LF, V, Uc, 2)
    assert np.abs(f.download() - V).max() < 1e-14
    M.coarsen_links(f, 2, c)
    assert np.abs(c.download() - Uc).max() < 1e-13
    e = M.Field(ctx, LF, LINK18, F64)
    plaq, cplaq = M.lattice_expansion(2, field(ctx, LC, Uc), e)
    np.testing.assert_array_equal(e.download(), orc.expand(LC, Uc, 2))
    assert abs(cplaq - orc.avg_plaquette(LC, Uc)) < 1e-14
    assert abs(plaq - orc.avg_plaquette(LF, orc.expand(LC, Uc, 2))) < 1e-14


def test_primitives_fp32(ctx, orc):
    U, Uc = orc.haar_links(LF, 4), orc.haar_links(LC, 8)
    f, c = field(ctx, LF, U, F32), M.Field(ctx, LC, LINK12, F32)
    U32 = f.download()
    M.coarsen_links(f, 2, c)
    assert np.abs(c.download() - orc.coarsen(LF, U32, 2)).max() < 1e-6
    c.upload(Uc)
    M.update_constrain(f, c, 2)
    V = U32.copy()
    orc.update_constrain(LF, V, c.download(), 2)
    assert np.abs(f.download() - V).max() < 1e-6


def test_checks(ctx):
    f = M.Field(ctx, LF, LINK18, F64).unit()
    with pytest.raises(M.MagError):
        M.coarsen_links(f, 3, M.Field(ctx, LC, LINK18, F64))
    with pytest.raises(M.MagError):
        M.coarsen_links(f, 2, M.Field(ctx, LC, LINK12, F32))
    assert M.check_constrained_plaquette(f, 2) == 1.0


def test_lattice_expansion_with_nersc_files(ctx, orc, tmp_path):
    Uc = orc.haar_links(LC, 12)
    src = tmp_path / "coarse.nersc"
    M.export_config_nersc(field(ctx, LC, Uc), src)
    coarse, fine = M.Field(ctx, LC, LINK18, F64), M.Field(ctx, LF, LINK18, F64)
    out = tmp_path / "expanded.nersc"
    plaq, cplaq = M.lattice_expansion(2, coarse, fine, src, out)
    assert abs(cplaq - orc.avg_plaquette(LC, Uc)) < 1e-14
    links, h = M.nersc_read(out)
    assert h["skip_third"] and h["dims"] == LF and h["ensemble_label"] == "lattice expansion"
    assert np.abs(links[:, :, :12] - orc.expand(LC, Uc, 2)[:, :, :12]).max() < 1e-15


def test_double_multigrid_keeps_the_constraint(ctx, orc):
    """Constrained fine trajectories keep the products of mag fine links;
    after the coarse trajectory and update_constrain they equal the new
    coarse links; the constrained plaquette is the coarse plaquette."""
    L = (8, 8, 8, 8)
    arg = M.Arg_chmc(mag=2, trajectory_length=4, dt=0.25, beta=5.7, num_forced_accept_step=100)
    arg_c = M.Arg_chmc(trajectory_length=4, dt=0.25, beta=5.0, num_forced_accept_step=100)
    U = orc.haar_links(L, 21)
    fine, coarse, rows = double_multigrid(ctx, U, L, arg, arg_c, num_trajectory=2, fine_per_coarse=2)
    Lc = (4, 4, 4, 4)
    Cf = orc.coarsen(L, fine.get_links(), 2)
    assert np.abs(Cf - coarse.get_links()).max() < 1e-12
    for i, cplaq, plaq_c, rf, rc in rows:
        assert abs(cplaq - plaq_c) < 1e-12
        assert all(np.isfinite(r.h_new) for r in rf) and np.isfinite(rc.h_new)
    # the fine trajectories alone leave the coarse products unchanged
    before = orc.coarsen(L, fine.get_links(), 2)
    fine.run(1)
    assert np.abs(orc.coarsen(L, fine.get_links(), 2) - before).max() < 1e-12
    assert abs(orc.avg_plaquette(Lc, before) - rows[-1][2]) < 1e-12
