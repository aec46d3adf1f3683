"""Padding sizes behind the C ABI (host code) against what the mathematics
fixes: FFTGrid::findClosestFactorableNumber (src/fftgrid.cpp:2837-2866) is
the smallest 2^a 3^b 5^c 7^d >= n, brute-forced here for every n up to
30000 and on 3000 random sizes up to 2^24, except where the reference's
loop bounds -- ceil of float logarithms -- fall one short: at n = 2^22 + 1,
2^22 + 2 and 2^24 + 1 below 2^24 + 2 it returns 2^22 (2^24), smaller than
n, and the library does the same.  setPaddingSize and estimate*PaddingSize
(src/modelgeneral.cpp:581-587,1337-1425) by hand-computed cases.
"""
import math

import numpy as np
import pytest

from crava_amd import _lib as L
from crava_amd import core


def _smooth7(limit):
    s = set()
    p2 = 1
    while p2 <= limit:
        p3 = p2
        while p3 <= limit:
            p5 = p3
            while p5 <= limit:
                p7 = p5
                while p7 <= limit:
                    s.add(p7)
                    p7 *= 7
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return np.array(sorted(s))


SMOOTH = _smooth7(2 ** 26)


def _brute(n):
    return int(SMOOTH[np.searchsorted(SMOOTH, n)])


def test_closest_factorable_brute_force_small():
    got = [core.padding_size(n, 0.0) for n in range(1, 30001)]
    want = [_brute(n) for n in range(1, 30001)]
    assert got == want


def test_closest_factorable_random_large_and_the_float_log_quirk():
    rng = np.random.default_rng(3)
    quirk = {2 ** 22 + 1: 2 ** 22, 2 ** 22 + 2: 2 ** 22, 2 ** 24 + 1: 2 ** 24}
    for n in list(rng.integers(30000, 2 ** 24, 3000)) + [2 ** 22, 2 ** 22 + 3, 2 ** 24]:
        assert core.padding_size(int(n), 0.0) == _brute(int(n)), n
    for n, v in quirk.items():
        assert core.padding_size(n, 0.0) == v


def test_set_padding_size():
    # ceil(n (1 + px)) with 1 + px formed from the float 1.0f
    assert core.padding_size(64, 0.1) == 72          # ceil(70.4) = 71 -> 72
    assert core.padding_size(256, 0.1) == 288        # ceil(281.6) = 282 -> 288
    assert core.padding_size(448, 0.1428) == 512     # ceil(511.97) = 512
    assert core.padding_size(100, 0.0) == 100
    assert core.padding_size(1, 1.0) == 2
    with pytest.raises(L.CravaHipError):
        core.padding_size(0, 0.1)


def test_estimate_xy_padding():
    # no estimate: the given factors
    nxp, nyp, fx, fy = core.estimate_xy_padding(100, 60, 25.0, 25.0, 2500.0, 1500.0, 0.1, 0.2, False, 0, 0, 0)
    assert (nxp, nyp) == (112, 72)                   # ceil(110) -> 112, ceil(72) -> 72
    assert fx == (112 - 100) / 100 and fy == (72 - 60) / 60
    # estimate from ranges: xPad = 0.5 max(|r1 cos a|, |r2 sin a|), at least dx, factor at most 1
    r1, r2, a = np.float32(800.0), np.float32(400.0), np.float32(0.0)
    nxp, nyp, fx, fy = core.estimate_xy_padding(100, 60, 25.0, 25.0, 2500.0, 1500.0, 0.0, 0.0, True, r1, r2, a)
    xpad, ypad = 0.5 * 800.0, 0.5 * 400.0
    assert nxp == core.padding_size(100, xpad / 2500.0) and nyp == core.padding_size(60, ypad / 1500.0)
    # ranges below one cell: one cell of padding
    nxp, nyp, _, _ = core.estimate_xy_padding(100, 60, 25.0, 25.0, 2500.0, 1500.0, 0.0, 0.0, True, 10.0, 10.0, 0.0)
    assert nxp == core.padding_size(100, 25.0 / 2500.0) and nyp == core.padding_size(60, 25.0 / 1500.0)
    # huge ranges: at most 100 % padding
    nxp, _, _, _ = core.estimate_xy_padding(100, 60, 25.0, 25.0, 2500.0, 1500.0, 0.0, 0.0, True, 1e6, 1e6, 0.3)
    assert nxp == core.padding_size(100, 1.0) == 200


def test_estimate_z_padding():
    nzp, fz = core.estimate_z_padding(80, 320.0, 0.5, 0.1, False, 0.0)
    assert nzp == 90 and fz == (90 - 80) / 80        # ceil(88) = 88 -> 90
    # default wavelet length 200 ms over lz * minRelThick = 160 ms: factor capped at 1
    nzp, fz = core.estimate_z_padding(80, 320.0, 0.5, 0.1, True, 200.0)
    assert nzp == 160 and fz == 1.0
    nzp, _ = core.estimate_z_padding(80, 320.0, 1.0, 0.1, True, 100.0)
    assert nzp == core.padding_size(80, 100.0 / 320.0) == 105
