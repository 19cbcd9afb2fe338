"""ILDG/LIME reading and writing on the host (ildg_io/ildgIo_gaugefield.cpp:44-131, ildg_io/checksum.hpp).

* The SciDAC checksum (CRC-32 per site, rotated by rank % 29 and rank % 31) of conf.00200's binary payload
  equals the checksum record the file itself carries (suma a3b09c8, sumb d7a6d9a1): pins the CRC, the
  rotation and the site order.
* conf.00200 read, written and read again gives bit-identical links, a file whose records parse in the
  reference's order with its checksum, and the oracle's plaquette 0.571077 of the re-read links.
* A 32-bit write reads back as the links rounded to float; a corrupted payload is rejected by the checksum;
  wrong extents are rejected.
"""
import os
import struct

import numpy as np
import pytest

import oracle as O
from cl2qcd_amd import _host as H
from conftest import CONF_00200


def records(path):
    d = open(path, "rb").read()
    off, out = 0, []
    while off + 144 <= len(d):
        magic, ver, flags, ln = struct.unpack(">IHHQ", d[off:off + 16])
        typ = d[off + 16:off + 144].split(b"\0")[0].decode()
        out.append((magic, ver, flags, typ, off + 144, ln))
        off += 144 + (ln + 7) // 8 * 8
    assert off == len(d)
    return d, out


def test_checksum_matches_the_files_own_record():
    d, recs = records(CONF_00200)
    (bin_off, bin_len), = [(o, n) for (_, _, _, t, o, n) in recs if t == "ildg-binary-data"]
    suma, sumb = H.ildg_checksum(d[bin_off:bin_off + bin_len], 4, 4)
    assert (suma, sumb) == (0xa3b09c8, 0xd7a6d9a1)
    _, info = H.ildg_read(CONF_00200, 4, 4)
    assert info["has_checksum"] and (info["suma"], info["sumb"]) == (suma, sumb)


def test_round_trip_of_conf_00200(tmp_path):
    links, _ = H.ildg_read(CONF_00200, 4, 4)
    lat = O.lattice(4, 4)
    assert abs(O.plaquette(lat, O.ildg_read(CONF_00200, lat)) - 0.571077) < 5e-7
    out = str(tmp_path / "conf.roundtrip")
    H.ildg_write(out, links, 4, 4, trajectory=200, plaquette=0.571077, beta=5.69, kappa=0.125)
    back, info = H.ildg_read(out, 4, 4)
    assert np.array_equal(back, links)
    assert info["precision"] == 64 and info["has_checksum"]
    assert "trajectory nr = 200" in info["xlf_info"] and "plaquette = 0.571077" in info["xlf_info"]
    d, recs = records(out)
    assert [t for (*_, t, _, _) in recs] == ["xlf-info", "ildg-format", "ildg-binary-data", "scidac-checksum"]
    assert all(m == 0x456789ab and v == 1 for (m, v, *_) in recs)
    assert [f for (_, _, f, *_) in recs] == [0xc000, 0x8000, 0x8000, 0x8000]
    # the same payload as the reference's file, byte for byte
    d0, recs0 = records(CONF_00200)
    (o0, n0), = [(o, n) for (*_, t, o, n) in recs0 if t == "ildg-binary-data"]
    (o1, n1), = [(o, n) for (*_, t, o, n) in recs if t == "ildg-binary-data"]
    assert d0[o0:o0 + n0] == d[o1:o1 + n1]
    # the oracle reads the written file and measures the same plaquette
    assert abs(O.plaquette(lat, O.ildg_read(out, lat)) - 0.571077) < 5e-7
    # writing over an existing file keeps the old one as a backup
    H.ildg_write(out, links, 4, 4, trajectory=201)
    assert os.path.exists(out + "_backup")


def test_single_precision_and_rejections(tmp_path):
    links, _ = H.ildg_read(CONF_00200, 4, 4)
    out = str(tmp_path / "conf.f32")
    H.ildg_write(out, links, 4, 4, precision=32)
    back, info = H.ildg_read(out, 4, 4)
    assert info["precision"] == 32
    assert np.array_equal(back, links.astype(np.float32).astype(np.float64))
    with pytest.raises(ValueError):
        H.ildg_read(out, 4, 8)  # extents differ from the file's
    d, recs = records(out)
    (o, n), = [(o, n) for (*_, t, o, n) in recs if t == "ildg-binary-data"]
    bad = bytearray(d)
    bad[o + 100] ^= 0x01
    corrupt = str(tmp_path / "conf.corrupt")
    open(corrupt, "wb").write(bytes(bad))
    with pytest.raises(RuntimeError):
        H.ildg_read(corrupt, 4, 4)
    back2, _ = H.ildg_read(corrupt, 4, 4, ignore_checksum=True)
    assert not np.array_equal(back2, back)
