"""MAOS .bin I/O (maos_amd/host/binio.c via maos_amd.telemetry)."""
import gzip
import os
import struct

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _build():
    from maos_amd import _lib
    _lib.setup()


def test_bin_bytes_follow_the_format(tmp_path):
    """Byte layout of one double array with keywords, built by hand from the
    format: M_COMMENT, length (rounded up to 8 with the NUL), text, length,
    M_COMMENT, M_SKIP, M_DBL, nx, ny, data."""
    from maos_amd import telemetry
    fn = str(tmp_path / "a.bin")
    a = np.arange(6, dtype=np.float64).reshape(2, 3)  # ny=2, nx=3
    telemetry.write_bin(fn, a, keywords="dx=0.5")
    raw = open(fn, "rb").read()
    txt = b"dx=0.5" + b"\0" * 2  # 6 chars + NUL -> 8
    want = (struct.pack("<IQ", 0x6500, 8) + txt + struct.pack("<QI", 8, 0x6500)
            + struct.pack("<IIQQ", 0x6600, 0x6402, 3, 2) + a.tobytes())
    assert raw == want


def test_bin_round_trip_plain_and_gz(tmp_path):
    from maos_amd import telemetry
    rs = np.random.RandomState(3)
    for name in ("x.bin", "x.bin.gz"):
        fn = str(tmp_path / name)
        for a in (rs.standard_normal((5, 7)), rs.standard_normal(11).astype(np.float32),
                  rs.standard_normal((2, 3)) + 1j * rs.standard_normal((2, 3))):
            telemetry.write_bin(fn, a, keywords="k=1\n")
            b, kw = telemetry.read_bin(fn, with_keywords=True)
            assert np.array_equal(b.reshape(a.shape), a) and kw.startswith("k=1")
    assert open(str(tmp_path / "x.bin.gz"), "rb").read(2) == b"\x1f\x8b"


def test_zfarr_cells_and_blanks(tmp_path):
    """A zfarr of 5 steps with steps 1 and 3 pushed reads back as a 5-cell
    with empty elements elsewhere (math/zfarr.c:57-110)."""
    from maos_amd import telemetry
    fn = str(tmp_path / "wfs0_gradcl_1.bin")
    z = telemetry.ZfArr(fn, 5)
    z.push(1, np.array([1.0, 2.0]))
    z.push(3, np.array([3.0, 4.0, 5.0]))
    z.close()
    cells = telemetry.read_bin(fn)
    assert len(cells) == 5
    assert cells[0] is None and cells[2] is None and cells[4] is None
    assert np.array_equal(cells[1].ravel(), [1, 2]) and np.array_equal(cells[3].ravel(), [3, 4, 5])


def test_reads_maos_sodium_profile():
    """The reference's NapGaussian.bin (gzip, written by MAOS): 1431 x 2
    doubles, heights at 24 m steps and a profile normalised to 1 whose mean
    is 90 km; parsed independently with gzip/struct for comparison."""
    from maos_amd import telemetry
    fn = os.path.join(GOLDEN, "NapGaussian.bin")
    a, kw = telemetry.read_bin(fn, with_keywords=True)
    assert a.shape == (2, 1431) and "Created by" in kw
    h, p = a
    assert np.allclose(np.diff(h), 24.0) and abs(p.sum() - 1) < 1e-12
    assert abs((h * p).sum() - 90e3) < 1.0
    raw = gzip.open(fn).read()
    off = raw.index(struct.pack("<I", 0x6402)) + 4 + 16
    ref = np.frombuffer(raw, dtype="<f8", count=2 * 1431, offset=off).reshape(2, 1431)
    assert np.array_equal(a, ref)


def test_sodium_profile_from_maos_file():
    """SodiumCfg(fnprof=...) feeds a MAOS profile file to the ETF builder in
    the layout of the synthetic Gaussian (nh x 2, column-major)."""
    from maos_amd.powfs import SodiumCfg
    t = SodiumCfg(fnprof=os.path.join(GOLDEN, "NapGaussian.bin")).table()
    nh = t.size // 2
    h, p = t[:nh], t[nh:]
    assert nh == 1431 and np.all(np.diff(h) > 0) and abs(p.sum() - 1) < 1e-12
    g = SodiumCfg().table()
    assert g.size % 2 == 0 and abs(g[g.size // 2:].sum() - 1) < 1e-12
