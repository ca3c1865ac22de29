"""NERSC configuration files (export_config_nersc / import_config_nersc,
field-matrix.h:445-713), host side: the writer and reader of the C ABI
checked against an independent numpy/zlib reading of the same bytes. No
device is touched here (the field-level calls are in test_gpu_nersc.py)."""
import zlib

import numpy as np
import pytest

import mag_amd as M
from mag_amd import _lib

L = (4, 4, 4, 8)


def split(path):
    raw = open(path, "rb").read()
    end = raw.index(b"END_HEADER\n") + len(b"END_HEADER\n")
    return raw[:end].decode().splitlines(), raw[end:]


@pytest.fixture
def links(orc):
    return orc.haar_links(L, 5)


@pytest.mark.parametrize("skip", [False, True])
def test_layout_checksum_crc_and_header(tmp_path, links, skip):
    p = tmp_path / "cfg.nersc"
    ck, crc = M.nersc_write(p, links, L, skip_third=skip, plaquette=0.123456789012345, link_trace=-0.00123456789,
                            beta=2.6, sequence_number=7, ensemble_label="test-ens")
    head, data = split(p)
    per = 12 if skip else 18
    # data: site order, mu, row-major (re, im), IEEE64 big-endian
    assert len(data) == np.prod(L) * 4 * per * 8
    np.testing.assert_array_equal(np.frombuffer(data, ">f8"), links[:, :, :per].ravel())
    # CHECKSUM: 32-bit sum of the data as 32-bit words; CRC32HASH: zlib crc32 of the data bytes
    assert ck == int(np.frombuffer(data, ">u4").astype(np.uint64).sum() % (1 << 32))
    assert crc == zlib.crc32(data)
    kv = dict(line.split(" = ", 1) for line in head[1:-1])
    assert head[0] == "BEGIN_HEADER" and head[-1] == "END_HEADER"
    assert list(kv)[:9] == ["HDR_VERSION", "DATATYPE", "DIMENSION_1", "DIMENSION_2", "DIMENSION_3", "DIMENSION_4",
                            "CRC32HASH", "CHECKSUM", "LINK_TRACE"]
    assert kv["DATATYPE"] == ("4D_SU3_GAUGE" if skip else "4D_SU3_GAUGE_3x3")
    assert [int(kv[f"DIMENSION_{i}"]) for i in range(1, 5)] == list(L)
    assert int(kv["CHECKSUM"], 16) == ck and int(kv["CRC32HASH"], 16) == crc
    prec = 12 if skip else 6  # ostream precision of the reference's two header variants
    assert kv["PLAQUETTE"] == format(0.123456789012345, f".{prec}g")
    assert kv["LINK_TRACE"] == format(-0.00123456789, f".{prec}g")
    assert kv["BETA"] == format(2.6, f".{prec}g")
    assert kv["FLOATING_POINT"] == "IEEE64BIG" and kv["SEQUENCE_NUMBER"] == "7"
    assert kv["ENSEMBLE_LABEL"] == "test-ens"
    # read back
    back, h = M.nersc_read(p)
    assert h["dims"] == L and h["skip_third"] == skip and h["checksum"] == ck and h["crc32"] == crc
    assert h["sequence_number"] == 7 and h["ensemble_label"] == "test-ens"
    np.testing.assert_array_equal(back[:, :, :per], links[:, :, :per])
    if skip:
        assert not back[:, :, 12:].any()


def test_read_errors(tmp_path, links):
    p = tmp_path / "cfg.nersc"
    M.nersc_write(p, links, L)
    raw = bytearray(open(p, "rb").read())
    bad = tmp_path / "bad.nersc"
    raw2 = raw.copy()
    raw2[-5] ^= 0x10  # one data bit
    bad.write_bytes(bytes(raw2))
    with pytest.raises(M.MagError) as e:
        M.nersc_read(bad)
    assert e.value.code == _lib.ERR_CHECKSUM
    bad.write_bytes(bytes(raw[:-8]))  # short data
    with pytest.raises(M.MagError) as e:
        M.nersc_read(bad)
    assert e.value.code == _lib.ERR_IO
    bad.write_bytes(bytes(raw).replace(b"4D_SU3_GAUGE_3x3", b"4D_SU3_GAUGE_4x4"))
    with pytest.raises(M.MagError) as e:
        M.nersc_read(bad, header_only=True)
    assert e.value.code == _lib.ERR_IO
    with pytest.raises(M.MagError) as e:
        M.nersc_read(tmp_path / "missing.nersc")
    assert e.value.code == _lib.ERR_IO
