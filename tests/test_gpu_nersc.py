"""NERSC export/import of device fields (field-matrix.h:452-713): round trips
of FP64 and FP32 link fields in both formats, header plaquette and link trace
against the oracle, and the import checks (lattice size, plaquette)."""
import numpy as np
import pytest

import mag_amd as M
from mag_amd import F32, F64, LINK12, LINK18, _lib

pytestmark = pytest.mark.gpu

L = (8, 8, 8, 4)


@pytest.mark.parametrize("skip", [False, True])
def test_export_import_round_trip_fp64(ctx, orc, tmp_path, skip):
    U = orc.haar_links(L, 31)
    src = M.Field(ctx, L, LINK18, F64).upload(U)
    p = tmp_path / "cfg"
    M.export_config_nersc(src, p, M.Arg_export(beta=6.0, sequence_num=3, ensemble_label="r"), does_skip_third=skip)
    _, h = M.nersc_read(p, header_only=True)
    prec = 12 if skip else 6
    assert h["plaquette"] == float(format(orc.avg_plaquette(L, U), f".{prec}g"))
    tr = sum(U[:, :, k].sum() for k in (0, 8, 16)) / (12 * np.prod(L))
    assert h["link_trace"] == float(format(tr, f".{prec}g"))
    dst = M.Field(ctx, L, LINK18, F64)
    h2 = M.import_config_nersc(dst, p)
    assert h2["sequence_number"] == 3 and h2["skip_third"] == skip
    back = dst.download()
    if skip:  # third row rebuilt by reunitarize: the Haar links are unitary to ~1e-14
        assert np.abs(back - U).max() < 1e-13
    else:
        np.testing.assert_array_equal(back, U)


def test_fp32_field_round_trip_and_checks(ctx, orc, tmp_path):
    U32 = M.Field(ctx, L, LINK12, F32).haar(9)
    p = tmp_path / "cfg32"
    M.export_config_nersc(U32, p)
    dst = M.Field(ctx, L, LINK18, F64)
    M.import_config_nersc(dst, p)
    np.testing.assert_array_equal(dst.download(), U32.download())
    # wrong lattice size
    with pytest.raises(M.MagError) as e:
        M.import_config_nersc(M.Field(ctx, (8, 8, 8, 8), LINK18, F64), p)
    assert e.value.code == _lib.ERR_GEOMETRY
    # header plaquette off by more than 1e-4
    links, h = M.nersc_read(p)
    q = tmp_path / "cfg_bad_plaq"
    M.nersc_write(q, links, L, plaquette=h["plaquette"] + 2e-4)
    with pytest.raises(M.MagError) as e:
        M.import_config_nersc(dst, q)
    assert e.value.code == _lib.ERR_CHECKSUM
