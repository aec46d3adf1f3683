"""The reference's NRLib SEG-Y writer as built in oracle/_ref (the checker of
the host mirror's FFTGrid::writeSegyFile / writeSgriFile), against what the
SEG-Y rev. 0 layout and the IBM floating-point format fix: the textual
header's card images, the binary header's fields, one 240-byte header and
nz IBM floats per defined trace in inline / crossline order, IBM values of
known numbers, and the SGRI header and big-endian body.  The STORM writer
(FFTGrid::writeStormFile on NRLib's ToString and WriteBinaryFloat) against
the format's fixed-6-decimal header fields, big-endian floats in k, j, i
order and the ascii rows without trailing blanks."""
import os
import struct

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.skipif(not oracle.ref_segy_available(), reason="oracle/_ref not built")


def ibm_to_float(w):
    if w & 0x7fffffff == 0:
        return 0.0
    sign = -1.0 if w >> 31 else 1.0
    return sign * (w & 0xffffff) / float(1 << 24) * 16.0 ** (((w >> 24) & 0x7f) - 64)


def _case(nx=4, ny=3, nz=5):
    rnxp = 2 * (nx // 2 + 1)
    g = np.zeros((nz + 2, ny + 1, rnxp), np.float32)
    for k in range(nz):
        g[k, :ny, :nx] = np.float32(1.0 + k) * np.array([[1.5, -118.625, 0.15625, 2.0 ** -60]] * ny, np.float32)[:, :nx]
    geom = dict(xmin=1000.0, ymin=2000.0, dx=25.0, dy=25.0, angle=0.0, lz=5 * 4.0, dz=4.0, zmin=100.0, zmax=120.0,
                il0=100.0, xl0=200.0, ilStepX=0.04, ilStepY=0.0, xlStepX=0.0, xlStepY=0.04)
    top = np.full((ny, nx), 100.0)
    bot = top + 20.0
    return g, geom, top, bot


def test_segy_layout_and_known_values(tmp_path):
    g, geom, top, bot = _case()
    path = str(tmp_path / "t")
    oracle.write_segy_ref(path, g, 4, 3, 5, geom, top, bot, 100.0, 0)
    raw = open(path + ".segy", "rb").read()
    text = raw[:3200].decode("cp037")
    cards = [text[80 * i:80 * i + 80] for i in range(40)]
    assert cards[0].startswith("C 1") and cards[39].startswith("C40 END EBCDIC")
    assert "THE CRAVA PROGRAM" in cards[13]
    bh = struct.unpack(">3i24h", raw[3200:3260])
    assert bh[:3] == (9999, 1, 1) and bh[5] == 4000 and bh[7] == 5 and bh[9] == 1   # dt us, samples, IBM format
    n = 4 * 3
    assert len(raw) == 3600 + n * (240 + 4 * 5)
    ils = []
    for t in range(n):
        h = raw[3600 + t * 260:3600 + t * 260 + 240]
        il, xl = struct.unpack(">i", h[8:12])[0], struct.unpack(">i", h[20:24])[0]    # SeisWorks 9, 21
        x, y = struct.unpack(">i", h[72:76])[0], struct.unpack(">i", h[76:80])[0]     # 73, 77
        assert struct.unpack(">hh", h[114:118]) == (5, 4000)
        ils.append((il, xl))
        i, j = round((x + 0.5 - 1000.0) / 25.0 - 0.5), round((y + 0.5 - 2000.0) / 25.0 - 0.5)
        assert (il, xl) == (int(0.5 + 100 + (x + 0.5 - 1000.0) * 0.04), int(0.5 + 200 + (y + 0.5 - 2000.0) * 0.04))
        w = struct.unpack(">5I", raw[3600 + t * 260 + 240:3600 + (t + 1) * 260])
        vals = [ibm_to_float(v) for v in w]
        assert np.allclose(vals, g[:5, j, i].astype(np.float64), rtol=1e-6, atol=0), (t, vals)
    assert ils == sorted(ils)
    # IBM words of known numbers (IBM System/360 single precision)
    w0 = struct.unpack(">I", raw[3840:3844])[0]
    assert w0 == 0x41180000                      # 1.5 = 0x18/16^1 ... (0.09375 * 16)
    vals = {g[0, 0, i]: None for i in range(4)}
    words = {}
    for t in range(n):
        h = raw[3600 + t * 260:3600 + t * 260 + 240]
        x = struct.unpack(">i", h[72:76])[0]
        words[round((x + 0.5 - 1000.0) / 25.0 - 0.5)] = struct.unpack(">I", raw[3600 + t * 260 + 240:3600 + t * 260 + 244])[0]
    assert words[1] == 0xC276A000                # -118.625
    assert words[2] == 0x40280000                # 0.15625


def test_segy_trace_without_top_is_left_out_and_formats_move_fields(tmp_path):
    g, geom, top, bot = _case()
    top[1, 2] = -99999.0
    for thf, (il_off, xl_off, x_off) in {0: (8, 20, 72), 1: (220, 20, 72), 2: (188, 192, 180), 3: (4, 20, 72),
                                         4: (180, 184, 72)}.items():
        path = str(tmp_path / f"f{thf}")
        oracle.write_segy_ref(path, g, 4, 3, 5, geom, top, bot, 100.0, thf)
        raw = open(path + ".segy", "rb").read()
        assert len(raw) == 3600 + 11 * 260
        h = raw[3600:3840]
        il = struct.unpack(">i", h[il_off:il_off + 4])[0]
        xl = struct.unpack(">i", h[xl_off:xl_off + 4])[0]
        x = struct.unpack(">i", h[x_off:x_off + 4])[0]
        assert (il, xl, x) == (101, 201, 1012), thf


def test_sgri_header_and_body(tmp_path):
    g, geom, top, bot = _case()
    path = str(tmp_path / "s")
    oracle.write_sgri_ref(path, "Label text", g, 4, 3, 5, geom, top, bot)
    lines = open(path + ".Sgrh").read().splitlines()
    assert lines[:9] == ["NORSAR General Grid Format v1.0", "3", "X (km)", "Y (km)", "T (s)", "FFT-grid", "1",
                         "Label text", "1 1 1"]
    assert lines[9] == "4 3 5"
    assert lines[10] == "0.025 0.025 0.004"
    assert lines[13] == "-99999" and lines[14] == path + ".Sgri" and lines[15] == "0"
    body = np.frombuffer(open(path + ".Sgri", "rb").read(), dtype=">f4").reshape(5, 3, 4)
    # z = zmin + k dz lies inside [top, bot]; sample k of the volume at k
    assert np.array_equal(body.astype(np.float32), g[:5, :3, :4])


def test_storm_binary_ascii_and_padded(tmp_path):
    """Header fields of Simbox::getStormHeader (ToString(x, 6) is fixed with 6
    decimals, the cube type and sizes plain integers, the angle in degrees),
    the cube in k, j, i order as big-endian floats or "%.6f" rows, then "0"."""
    nx, ny, nz = 5, 3, 4
    rnxp, nyp, nzp = 2 * (6 // 2 + 1), 4, 6
    rng = np.random.default_rng(3)
    g = rng.standard_normal((nzp, nyp, rnxp)).astype(np.float32) * np.float32(100.0)
    g[0, 0, 0] = np.float32(2.0 ** -20)
    head = dict(xmin=456000.5, lx=11200.0, ymin=6780000.25, ly=5600.0, lz=1024.0, angle=0.3,
                top="top_surface", bot="base_surface")
    f6 = lambda v: "%.6f" % v
    geo = f"{f6(456000.5)} {f6(11200.0)} {f6(6780000.25)} {f6(5600.0)}"

    def storm_head(kind, geom, n3):
        return (f"storm_petro_{kind}\n0 4 {f6(-99999.0)}\nFFTGrid\n{geom}\n"
                f"{f6(1024.0)} {f6(0.3 * 180 / np.pi)}\n\n{n3[0]} {n3[1]} {n3[2]}\n")
    oracle.write_storm_ref(str(tmp_path / "a"), g, nx, ny, nz, 4, head)
    want = storm_head("binary", geo + " top_surface base_surface 0.0 0.0", (nx, ny, nz)).encode()
    want += g[:nz, :ny, :nx].astype(">f4").tobytes() + b"0\n"
    assert open(tmp_path / "a.storm", "rb").read() == want
    oracle.write_storm_ref(str(tmp_path / "p"), g, 6, nyp, nzp, 4, head)
    want = storm_head("binary", geo + " top_surface base_surface 0.0 0.0", (6, nyp, nzp)).encode()
    want += g[:, :, :6].astype(">f4").tobytes() + b"0\n"
    assert open(tmp_path / "p.storm", "rb").read() == want
    oracle.write_storm_ref(str(tmp_path / "f"), g, nx, ny, nz, 4, head, ascii=True, flat=True)
    lines = ["%s\n" % " ".join(f6(float(x)) for x in g[k, j, :nx]) for k in range(nz) for j in range(ny)]
    want = storm_head("ascii", geo + f" 0.0 {f6(1024.0)} 0.0 0.0", (nx, ny, nz)) + "".join(lines) + "0\n"
    assert open(tmp_path / "f.txt").read() == want
