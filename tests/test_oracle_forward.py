"""The oracle's forward modelling (Crava::computeSyntSeismic,
src/crava.cpp:1667-1759) pinned without retyping it:

* against the reference's own FFTW 2.1.2 driving the same trace loop
  (oracle/_ref, when built) and the golden vectors it wrote
  (tests/golden/forward_cases.npz);
* closed forms of the filter: a delta wavelet at sample m returns the
  reflection coefficients advanced by m; an integer local time shift s
  advances them by s more; a gain scales them; a stretch sf >= 1 divides
  them by sf, and sf = 1/2 keeps the bins below nzp/4 and doubles them;
* brute force: a random wavelet equals the circular cross-correlation of
  the reflection coefficients with it, in float64;
* the taper and difference: a linear impedance ramp gives its slope inside
  the volume, zero where the taper meets the volume's ends, and the taper's
  slope in between;
* padding traces hold the impedance, bit for bit.
"""
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "forward_cases.npz")


def _case(seed=0, nx=7, ny=5, nz=20, nxp=9, nyp=6, nzp=32, nt=2):
    rng = np.random.default_rng(seed)
    rnxp = 2 * (nxp // 2 + 1)
    g = [rng.standard_normal((nzp, nyp, rnxp)).astype(np.float32) for _ in range(3)]
    A = rng.standard_normal((nt, 3)).astype(np.float32)
    return (nx, ny, nz, nxp, nyp, nzp), g, A


def _delta(nt, nzp, m):
    w = np.zeros((nt, nzp), np.float32)
    w[:, m % nzp] = 1.0
    return w


def test_delta_wavelet_advances_the_reflection_coefficients():
    n, g, A = _case()
    nx, ny, nz, nxp, nyp, nzp = n
    base = oracle.synthetic_seismic(*n, A, _delta(2, nzp, 0), 4.0, *g)
    for m in (1, 5, nzp - 3):
        out = oracle.synthetic_seismic(*n, A, _delta(2, nzp, m), 4.0, *g)
        for l in range(2):
            want = np.roll(base[l][:, :ny, :nx], -m, axis=0)
            assert np.allclose(out[l][:, :ny, :nx], want, atol=2e-6 * np.abs(want).max()), (m, l)


def test_local_shift_gain_and_stretch_closed_forms():
    n, g, A = _case(seed=1)
    nx, ny, nz, nxp, nyp, nzp = n
    dz = 4.0
    w = _delta(2, nzp, 0)
    base = oracle.synthetic_seismic(*n, A, w, dz, *g)
    rng = np.random.default_rng(5)
    s_int = rng.integers(-6, 7, size=(ny, nx))
    gain = rng.uniform(0.5, 2.0, size=(ny, nx)).astype(np.float32)
    out = oracle.synthetic_seismic(*n, A, w, dz, *g, shift=[(s_int * dz).astype(np.float32), None],
                                   gain=[gain, gain])
    for j in range(ny):
        for i in range(nx):
            want0 = gain[j, i] * np.roll(base[0][:, j, i], -int(s_int[j, i]))
            assert np.allclose(out[0][:, j, i], want0, atol=3e-6 * np.abs(base[0]).max()), (i, j)
            assert np.allclose(out[1][:, j, i], gain[j, i] * base[1][:, j, i], atol=3e-6 * np.abs(base[1]).max())
    # stretch
    two = np.full((ny, nx), 2.0, np.float32)
    half = np.full((ny, nx), 0.5, np.float32)
    o2 = oracle.synthetic_seismic(*n, A, w, dz, *g, stretch=[two, half])
    assert np.allclose(o2[0][:, :ny, :nx], base[0][:, :ny, :nx] / 2, atol=2e-6 * np.abs(base[0]).max())
    spec = np.fft.rfft(base[1][:, :ny, :nx].astype(np.float64), axis=0)
    k = np.arange(nzp // 2 + 1)
    spec[2 * k >= nzp // 2 + 1] = 0.0
    want = 2.0 * np.fft.irfft(spec, n=nzp, axis=0)
    assert np.allclose(o2[1][:, :ny, :nx], want, atol=5e-6 * np.abs(want).max())


def test_random_wavelet_is_a_circular_cross_correlation():
    n, g, A = _case(seed=2)
    nx, ny, nz, nxp, nyp, nzp = n
    rng = np.random.default_rng(3)
    w = rng.standard_normal((2, nzp)).astype(np.float32)
    r = oracle.synthetic_seismic(*n, A, _delta(2, nzp, 0), 4.0, *g)
    out = oracle.synthetic_seismic(*n, A, w, 4.0, *g)
    for l in range(2):
        rr = r[l][:, :ny, :nx].astype(np.float64)
        # out(t) = sum_m r(t + m) w(m)
        want = sum(np.roll(rr, -m, axis=0) * float(w[l, m]) for m in range(nzp))
        assert np.allclose(out[l][:, :ny, :nx], want, atol=5e-6 * np.abs(want).max()), l


def test_linear_impedance_ramp_taper_and_difference():
    nx, ny, nz, nxp, nyp, nzp = 4, 3, 24, 5, 4, 40
    rnxp = 2 * (nxp // 2 + 1)
    a0, b = 0.75, 0.125
    k = np.arange(nzp, dtype=np.float32)[:, None, None]
    alpha = np.broadcast_to(np.float32(a0) + np.float32(b) * k, (nzp, nyp, rnxp)).astype(np.float32)
    zero = np.zeros_like(alpha)
    A = np.array([[1.0, 0.0, 0.0]], np.float32)
    out = oracle.synthetic_seismic(nx, ny, nz, nxp, nyp, nzp, A, _delta(1, nzp, 0), 4.0, alpha, zero, zero)[0]
    tr = out[:, 1, 2].astype(np.float64)
    first, last = a0, a0 + b * (nz - 1)
    assert np.allclose(tr[: nz - 1], b, atol=1e-5)
    assert abs(tr[nz - 1]) < 1e-5 and abs(tr[nzp - 1]) < 1e-5
    assert np.allclose(tr[nz: nzp - 1], (first - last) / (nzp - nz - 1), atol=1e-5)
    assert abs(tr.sum()) < 1e-4


def test_padding_traces_hold_the_impedance():
    n, g, A = _case(seed=4)
    nx, ny, nz, nxp, nyp, nzp = n
    rng = np.random.default_rng(9)
    out = oracle.synthetic_seismic(*n, A, rng.standard_normal((2, nzp)).astype(np.float32), 4.0, *g)
    for l in range(2):
        imp = np.float32(0) + g[0] * A[l, 0]
        imp = imp + g[1] * A[l, 1]
        imp = imp + g[2] * A[l, 2]
        mask = np.ones(g[0].shape[1:], bool)
        mask[:ny, :nx] = False
        assert np.array_equal(out[l][:, mask], imp[:, mask])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs the reference sources)")
def test_oracle_matches_reference_fftw_path():
    n, g, A = _case(seed=6, nt=3)
    nx, ny, nz, nxp, nyp, nzp = n
    rng = np.random.default_rng(7)
    w = rng.standard_normal((3, nzp)).astype(np.float32)
    sh = [rng.uniform(-8, 8, (ny, nx)).astype(np.float32), None, rng.uniform(-3, 3, (ny, nx)).astype(np.float32)]
    gn = [None, rng.uniform(0.5, 1.5, (ny, nx)).astype(np.float32), None]
    st = [rng.uniform(0.8, 1.25, (ny, nx)).astype(np.float32), None, None]
    kw = dict(shift=sh, gain=gn, stretch=st)
    o = oracle.synthetic_seismic(*n, A, w, 4.0, *g, **kw)
    r = oracle.synthetic_seismic(*n, A, w, 4.0, *g, which="ref", **kw)
    for a, b in zip(o, r):
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


def test_oracle_matches_golden_vectors():
    d = np.load(GOLDEN)
    nx, ny, nz, nxp, nyp, nzp = (int(x) for x in d["dims"])
    nt = int(d["A"].shape[0])
    tabs = {}
    for name in ("shift", "gain", "stretch"):
        tabs[name] = [d[f"{name}{l}"] if f"{name}{l}" in d else None for l in range(nt)]
    out = oracle.synthetic_seismic(nx, ny, nz, nxp, nyp, nzp, d["A"], d["wavelets"], float(d["dz"]), d["alpha"],
                                   d["beta"], d["rho"], **tabs)
    for l in range(nt):
        want = d[f"out{l}"]
        assert np.abs(out[l] - want).max() <= 1e-6 * np.abs(want).max(), l
