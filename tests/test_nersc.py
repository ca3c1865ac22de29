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
    # CHECKSUM: 32-bit sum of the host-order doubles as 32-bit words (each big-endian word of the
    # stored data is one half of a double); CRC32HASH: zlib crc32 of the data in host byte order
    assert ck == int(np.frombuffer(data, ">u4").astype(np.uint64).sum() % (1 << 32))
    assert crc == zlib.crc32(np.frombuffer(data, ">f8").astype("<f8").tobytes())
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
    assert not h["checksum_swapped"]
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


def reference_two_row_file(path, links, dims):
    """A 4D_SU3_GAUGE file laid out as the reference's export writes it
    (field-matrix.h:466-517): the first two rows of every link as IEEE64
    big-endian, CHECKSUM the 32-bit sum taken after the byte swap, i.e. over
    the stored bytes read as little-endian 32-bit words."""
    data = np.ascontiguousarray(links[:, :, :12]).astype(">f8").tobytes()
    ck = int(np.frombuffer(data, "<u4").astype(np.uint64).sum() % (1 << 32))
    head = ["BEGIN_HEADER", "HDR_VERSION = 1.0", "DATATYPE = 4D_SU3_GAUGE"]
    head += [f"DIMENSION_{i + 1} = {d}" for i, d in enumerate(dims)]
    head += ["CRC32HASH = 0", f"CHECKSUM = {ck:x}", "LINK_TRACE = 0", "PLAQUETTE = 0", "CREATOR = RBC",
             "ARCHIVE_DATE = Thu Jan  1 00:00:00 1970", "ENSEMBLE_LABEL = ref", "FLOATING_POINT = IEEE64BIG",
             "ENSEMBLE_ID = NOT yet implemented", "SEQUENCE_NUMBER = 1", "BETA = 2.6", "END_HEADER"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        f.write(data)
    return ck


def test_reads_two_row_file_with_the_reference_export_checksum(tmp_path, links):
    p = tmp_path / "ref12.nersc"
    ck = reference_two_row_file(p, links, L)
    host_order = int(np.frombuffer(links[:, :, :12].tobytes(), "<u4").astype(np.uint64).sum() % (1 << 32))
    assert ck != host_order  # the two sums differ on this data, so the test tells them apart
    back, h = M.nersc_read(p)
    assert h["skip_third"] and h["checksum"] == ck and h["checksum_swapped"]
    np.testing.assert_array_equal(back[:, :, :12], links[:, :, :12])
    # the 3x3 format keeps the reference's single (host-order) rule
    p3 = tmp_path / "ref18.nersc"
    raw = open(p, "rb").read().replace(b"4D_SU3_GAUGE\n", b"4D_SU3_GAUGE_3x3\n")
    end = raw.index(b"END_HEADER\n") + len(b"END_HEADER\n")
    d18 = np.ascontiguousarray(links).astype(">f8").tobytes()
    ck18 = int(np.frombuffer(d18, "<u4").astype(np.uint64).sum() % (1 << 32))
    head = raw[:end].replace(f"CHECKSUM = {ck:x}".encode(), f"CHECKSUM = {ck18:x}".encode())
    p3.write_bytes(head + d18)
    with pytest.raises(M.MagError) as e:
        M.nersc_read(p3)
    assert e.value.code == _lib.ERR_CHECKSUM
