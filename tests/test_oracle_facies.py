"""Pins the oracle's facies density lookup and probability assembly
(orc_facies / orc_facies_find_density in oracle/crava_oracle.c) without
restating its loops.

The reference: PosteriorElasticPDF3D::FindDensity
(src/posteriorelasticpdf3d.cpp:511-604) with Simbox::getInterpolationIndexes
(src/simbox.cpp:186-201), FaciesProb::FindDensityFromPosteriorPDF
(src/faciesprob.cpp:1277-1317) and the grid loop of
CalculateFaciesProbFromPosteriorElasticPDF (src/faciesprob.cpp:2198-2265).

* brute_density is a second, independent statement of the lookup: clip each
  fractional index to [0, n-1] first, then interpolate linearly between
  floor and floor+1 in float64 (the reference branches on floor(index)
  instead and forms float weights);
* hand-computed cases with dyadic coordinates, where the lookup is exact: a
  table node returns the node, a cell centre the mean of its 8 corners,
  indices below 0 / above n-1 take the edge with zero weight, negative table
  entries count as 0;
* the probability grids: prior * density / (undefSum + sum), the undefined
  class undefSum / sum, RMISSING on the padding columns nx <= k < 2(nx/2+1);
* a mutation check: the brute force run on a table with two axes swapped
  disagrees with the oracle, so the comparison sees an axis mistake.
"""
import numpy as np
import pytest

import oracle

RMISSING = np.float32(-99999.0)


def _vol(x0, y0, dx, dy, z0, z1):
    return np.array([x0, y0, dx, dy, z0, z1], np.float64)


def brute_density(a, b, r, table, vol, vn):
    """Trilinear lookup of max(table, 0) at (a, b, r), clip-then-interpolate, float64."""
    n1, n2, n3 = (int(v) for v in vn)
    t = np.maximum(np.asarray(table, np.float64)[:, :, :n1], 0.0)   # [n3][n2][n1]
    a, b, r = (np.asarray(v, np.float64) for v in (a, b, r))
    fx = (a - vol[0]) / vol[2] - 0.5
    fy = (b - vol[1]) / vol[3] - 0.5
    fz = n3 * (r - vol[4]) / (vol[5] - vol[4]) - 0.5

    def axis(f, n):
        f = np.clip(f, 0.0, n - 1.0)
        lo = np.minimum(np.floor(f).astype(np.int64), n - 1)
        return lo, np.minimum(lo + 1, n - 1), f - lo

    x0, x1, wx = axis(fx, n1)
    y0, y1, wy = axis(fy, n2)
    z0, z1, wz = axis(fz, n3)
    out = np.zeros(np.broadcast(a, b, r).shape)
    for zi, zw in ((z0, 1 - wz), (z1, wz)):
        for yi, yw in ((y0, 1 - wy), (y1, wy)):
            for xi, xw in ((x0, 1 - wx), (x1, wx)):
                out = out + zw * yw * xw * t[zi, yi, xi]
    return out


def _table(seed, n1, n2, n3, neg=0.0):
    rng = np.random.default_rng(seed)
    rn = 2 * (n1 // 2 + 1)
    t = np.zeros((n3, n2, rn), np.float32)
    t[:, :, :n1] = rng.uniform(-neg, 1.0, (n3, n2, n1)).astype(np.float32)
    t[:, :, n1:] = 1e9   # padding columns must never be read
    return t


def _dens(a, b, r, tab, vol, vn):
    return oracle.facies_find_density(a, b, r, tab, (int(vn[0]), int(vn[1]), int(vn[2])), vol, vn)


# --------------------------------------------------------------- hand-computed
N = (6, 5, 4)
VOL = _vol(1.0, 2.0, 0.5, 0.25, 0.0, 2.0)     # dz = 0.5 on 4 cells


def _coord(j, k, l):
    """Elastic values whose fractional indices are exactly (j, k, l)."""
    return (VOL[0] + (j + 0.5) * VOL[2], VOL[1] + (k + 0.5) * VOL[3],
            VOL[4] + (l + 0.5) * (VOL[5] - VOL[4]) / N[2])


def test_node_returns_the_node_value():
    tab = _table(0, *N)
    vn = np.array(N, np.int32)
    for j, k, l in [(0, 0, 0), (2, 3, 1), (5, 4, 3), (4, 1, 2)]:
        assert _dens(*_coord(j, k, l), tab, VOL, vn) == float(tab[l, k, j])


def test_cell_centre_is_the_mean_of_its_eight_corners():
    tab = _table(1, *N)
    vn = np.array(N, np.int32)
    for j, k, l in [(0, 0, 0), (1, 2, 1), (4, 3, 2)]:
        got = _dens(*_coord(j + 0.5, k + 0.5, l + 0.5), tab, VOL, vn)
        want = float(np.mean(tab[l:l + 2, k:k + 2, j:j + 2].astype(np.float64)))
        assert got == pytest.approx(want, rel=1e-15, abs=0)


def test_indices_outside_the_table_take_the_edge():
    tab = _table(2, *N)
    vn = np.array(N, np.int32)
    # below 0 on x (index -0.7, -3): column 0 with zero weight
    for f in (-0.7, -3.0):
        assert _dens(*_coord(f, 2, 1), tab, VOL, vn) == float(tab[1, 2, 0])
    # at and above n-1 on y (index 4, 4.3, 9): row n2-1
    for f in (4.0, 4.3, 9.0):
        assert _dens(*_coord(3, f, 2), tab, VOL, vn) == float(tab[2, 4, 3])
    # above on z and below on x together
    assert _dens(*_coord(-1.5, 1, 7.2), tab, VOL, vn) == float(tab[3, 1, 0])
    # inside on x, clamped on z: linear in x only
    got = _dens(*_coord(2.25, 3, -2.0), tab, VOL, vn)
    want = 0.75 * float(tab[0, 3, 2]) + 0.25 * float(tab[0, 3, 3])
    assert got == pytest.approx(want, rel=1e-7)


def test_negative_entries_count_as_zero():
    tab = _table(3, *N)
    vn = np.array(N, np.int32)
    tab[1, 2, 3] = -0.4
    assert _dens(*_coord(3, 2, 1), tab, VOL, vn) == 0.0
    tab[1:3, 2:4, 3:5] = np.array([[[-1.0, 0.5], [0.25, -2.0]], [[1.0, -0.125], [-0.75, 0.5]]], np.float32)
    got = _dens(*_coord(3.5, 2.5, 1.5), tab, VOL, vn)
    assert got == pytest.approx((0.5 + 0.25 + 1.0 + 0.5) / 8, rel=1e-15)


# --------------------------------------------------------------- brute force
def test_lookup_matches_brute_force_on_random_points():
    n = (9, 7, 8)
    tab = _table(4, *n, neg=0.2)
    vol = _vol(-0.3, 0.7, 0.11, 0.07, 1.1, 1.9)
    vn = np.array(n, np.int32)
    rng = np.random.default_rng(5)
    # covers the table and a margin of two cells beyond every edge
    a = rng.uniform(vol[0] - 2 * vol[2], vol[0] + (n[0] + 2) * vol[2], 400)
    b = rng.uniform(vol[1] - 2 * vol[3], vol[1] + (n[1] + 2) * vol[3], 400)
    r = rng.uniform(vol[4] - 0.2, vol[5] + 0.2, 400)
    want = brute_density(a, b, r, tab, vol, vn)
    got = np.array([_dens(x, y, z, tab, vol, vn) for x, y, z in zip(a, b, r)])
    assert np.allclose(got, want, rtol=2e-6, atol=1e-7)


def test_swapped_table_axes_are_detected():
    """The brute-force comparison fails for a lookup with two axes swapped."""
    n = (8, 8, 6)
    tab = _table(6, *n)
    vol = _vol(0.0, 0.0, 0.1, 0.1, 0.0, 1.0)
    vn = np.array(n, np.int32)
    rng = np.random.default_rng(7)
    a, b, r = rng.uniform(0.05, 0.75, 100), rng.uniform(0.05, 0.75, 100), rng.uniform(0.05, 0.95, 100)
    got = np.array([_dens(x, y, z, tab, vol, vn) for x, y, z in zip(a, b, r)])
    assert np.allclose(got, brute_density(a, b, r, tab, vol, vn), rtol=2e-6, atol=1e-7)
    swapped = np.zeros_like(tab)
    swapped[:, :, :8] = np.swapaxes(tab[:, :, :8], 1, 2)
    assert not np.allclose(got, brute_density(a, b, r, swapped, vol, vn), rtol=1e-3)


# --------------------------------------------------------------- probabilities
def test_probability_grids_follow_the_assembly():
    nx, ny, nz, nxp, nyp, nzp = 7, 5, 4, 10, 6, 6
    tn = (9, 7, 8)
    nfac = 3
    rnxp = 2 * (nxp // 2 + 1)
    sm = 2 * (nx // 2 + 1)
    tabs = np.stack([_table(10 + f, *tn, neg=0.1) for f in range(nfac)])
    vol = _vol(-0.3, 0.7, 0.11, 0.07, 1.1, 1.9)
    vn = np.array(tn, np.int32)
    rng = np.random.default_rng(11)
    g = [rng.uniform(lo, hi, (nzp, nyp, rnxp)).astype(np.float32)
         for lo, hi in ((-0.4, 0.8), (0.6, 1.3), (1.0, 2.0))]
    prior = np.array([0.5, 0.3, 0.2], np.float32)
    p_undef = np.float32(0.02)
    out = oracle.facies(nx, ny, nz, nxp, nyp, nzp, *g, tabs, tn, vol, vn, prior, p_undef)

    undef_sum = np.float32(p_undef / np.float32(tn[0] * tn[1] * tn[2]))
    a, b, r = (x[:nz, :ny, :nx].astype(np.float64) for x in g)
    dens = np.stack([brute_density(a, b, r, tabs[f], vol, vn) for f in range(nfac)])
    val = prior[:, None, None, None].astype(np.float64) * dens
    tot = float(undef_sum) + val.sum(0)
    for f in range(nfac):
        assert np.allclose(out[f][:, :, :nx], val[f] / tot, rtol=1e-5, atol=1e-9)
    assert np.allclose(out[nfac][:, :, :nx], float(undef_sum) / tot, rtol=1e-5)
    for f in range(nfac + 1):
        assert out[f].shape == (nz, ny, sm)
        assert np.all(out[f][:, :, nx:] == RMISSING)
    # every voxel's classes sum to one
    s = sum(out[f][:, :, :nx].astype(np.float64) for f in range(nfac + 1))
    assert np.allclose(s, 1.0, atol=1e-6)


def test_zero_density_everywhere_gives_the_undefined_class():
    nx, ny, nz = 4, 3, 2
    tn = (5, 5, 5)
    tabs = np.zeros((2, 5, 5, 6), np.float32) - 1.0     # all negative -> density 0
    vol = _vol(0.0, 0.0, 0.2, 0.2, 0.0, 1.0)
    g = [np.full((nz, ny, 6), 0.5, np.float32) for _ in range(3)]
    out = oracle.facies(nx, ny, nz, nx, ny, nz, *g, tabs, tn, vol, np.array(tn, np.int32),
                        np.array([0.6, 0.4], np.float32), 0.1)
    assert np.all(out[0][:, :, :nx] == 0) and np.all(out[1][:, :, :nx] == 0)
    assert np.all(out[2][:, :, :nx] == 1.0)
