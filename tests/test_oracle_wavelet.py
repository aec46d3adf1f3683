"""The oracle's non-constant-thickness per-k tables, Wavelet::resample and
the energy-matching variances, pinned by what the mathematics fixes:

- per-k tables: kD is the spectrum of the forward-difference operator,
  -1 + e^{i w} (w = 2 pi k / nzp); kD3 that of its backward then central
  difference, kD (1 - e^{-i w}) i sin w; |errorSmooth3| by numpy's FFT of the
  same three differences; the robust inverse amplitude by numpy's rfft with
  the 0.999 stretch written as an interpolation, a delta wavelet (0.999
  everywhere; 1 where the stretch reads past Nyquist) and a zero wavelet
  (|w| nzp^2 100);
- resample: exact at the wavelet's own nodes, exact reproduction of a 20 Hz
  Ricker centred on a node inside the six-sample stencil (the simple-kriging
  predictor reproduces every covariance column), the Cholesky solve against
  numpy.linalg.solve, linearity and the norm;
- energy: a delta temporal correlation gives the sum of squared forward
  differences, a general correlation the Toeplitz quadratic form in float64,
  and matched angles end with theoretical S/N equal to the empirical one.
"""
import numpy as np
import pytest

import oracle


def _cx(t):
    return t[..., 0].astype(np.float64) + 1j * t[..., 1]


def _tables(w, wnorm, nzp, lz=400.0, nz=100):
    return oracle.perk_tables_nonconst(w.shape[0], nz, nzp, lz, 5.0, 55.0, 1.0, w, wnorm)


@pytest.mark.parametrize("nzp", [64, 90, 128, 75])
def test_difference_operator_spectra_closed_form(nzp):
    rng = np.random.default_rng(nzp)
    nt = 3
    w = rng.standard_normal((nt, nzp)).astype(np.float32)
    wnorm = np.float32(1.0) + rng.random(nt).astype(np.float32)
    t = _tables(w, wnorm, nzp)
    om = 2 * np.pi * np.arange(nzp) / nzp
    d1 = -1 + np.exp(1j * om)
    d3 = d1 * (1 - np.exp(-1j * om)) * 1j * np.sin(om)
    for l in range(nt):
        assert np.abs(_cx(t["kWD"])[:, l] - d1).max() < 2e-6
        assert np.array_equal(t["e1"][:, l, 0], np.full(nzp, np.float32(1.0) / wnorm[l]))
        assert not t["e1"][:, l, 1].any()
        # |errorSmooth3|: third difference of the wavelet by numpy's FFT
        c = np.fft.ifft(np.fft.fft(w[l].astype(np.float64)) * (np.exp(1j * om) - 1) * (1 - np.exp(-1j * om))
                        * 1j * np.sin(om)).real
        want = d3 / np.linalg.norm(c)
        got = _cx(t["e2"])[:, l]
        assert np.abs(got - want).max() < 3e-6 * np.abs(want).max()


@pytest.mark.parametrize("nzp", [64, 90, 128])
def test_robust_inverse_amplitude_vs_numpy(nzp):
    rng = np.random.default_rng(3 + nzp)
    nt = 2
    w = rng.standard_normal((nt, nzp)).astype(np.float32)
    t = _tables(w, np.ones(nt, np.float32), nzp)
    cn = nzp // 2 + 1
    for l in range(nt):
        W = np.fft.rfft(w[l].astype(np.float64))
        om = np.arange(cn) / 0.999
        lo = np.floor(om).astype(int)
        hi = np.minimum(lo + 1, cn - 1)
        d = om - lo
        Ws = np.where(om < cn, (W[np.minimum(lo, cn - 1)] * (1 - d) + W[hi] * d) / 0.999, 0)
        e3 = 1 / np.sqrt(np.maximum(np.abs(W) ** 2, np.abs(Ws) ** 2))
        full = np.concatenate([e3, e3[1:nzp - cn + 1][::-1]])
        got = t["e3"][:, l, 0].astype(np.float64)
        assert np.abs(got / full - 1).max() < 2e-5
        assert not t["e3"][:, l, 1].any()


def test_robust_inverse_amplitude_delta_and_zero_wavelets():
    nzp = 2048
    w = np.zeros((2, nzp), np.float32)
    w[0, 0] = 1.0
    wn = np.array([1.0, 2.5], np.float32)
    t = _tables(w, wn, nzp, lz=4000.0, nz=1000)
    e3 = t["e3"][:, 0, 0]
    # W = 1: the stretched spectrum is 1/0.999 while k/0.999 < nzp/2 + 1
    k = np.minimum(np.arange(nzp), nzp - np.arange(nzp))
    past = np.float32(k) / np.float32(0.999) >= nzp // 2 + 1
    assert past.sum() == 1                       # the Nyquist bin only, at nzp 2048
    np.testing.assert_allclose(e3[~past], 0.999, rtol=2e-7)
    assert np.all(e3[past] == 1.0)
    # zero wavelet: |w| nzp^2 100
    assert np.all(t["e3"][:, 1, 0] == np.float32(wn[1] * nzp * nzp * 100.0))


def test_frequency_gate_matches_constant_thickness_tables():
    rng = np.random.default_rng(5)
    nzp = 96
    w = rng.standard_normal((2, nzp)).astype(np.float32)
    wn = np.ones(2, np.float32)
    a = oracle.perk_tables(2, 80, nzp, 330.0, 5.0, 55.0, 0.7, w, wn)
    b = oracle.perk_tables_nonconst(2, 80, nzp, 330.0, 5.0, 55.0, 0.7, w, wn)
    assert np.array_equal(a["invert"], b["invert"])
    assert 0 < b["invert"].sum() < nzp


# ------------------------------------------------------------- resample
def _ricker(lag_ms):
    f = (20.0 * np.pi * lag_ms * 1e-3) ** 2
    return (1 - 2 * f) * np.exp(-f)


def _fft_order_z(dz, nzp):
    k = np.arange(nzp)
    return np.where(k < nzp // 2 + 1, np.float32(dz) * k.astype(np.float32),
                    np.float32(dz) * (k - nzp).astype(np.float32)).astype(np.float32)


def test_resample_exact_at_the_wavelets_own_nodes():
    rng = np.random.default_rng(11)
    nw, center, dz, nzp = 61, 30, 4.0, 80
    w = rng.standard_normal(nw).astype(np.float32)
    out, norm = oracle.wavelet_resample(w, center, dz, dz, nzp)
    k = np.arange(nzp)
    idx = np.where(k < nzp // 2 + 1, k, k - nzp) + center
    want = np.where((idx >= 0) & (idx < nw), w[np.clip(idx, 0, nw - 1)], 0).astype(np.float32)
    np.testing.assert_allclose(out, want, rtol=1e-6, atol=1e-6 * np.abs(w).max())
    assert norm == pytest.approx(float(np.sqrt(np.sum(out.astype(np.float64) ** 2))), rel=1e-6)


@pytest.mark.parametrize("dzw,dz", [(4.0, 1.3), (4.0, 2.5), (2.0, 0.7), (3.0, 4.0)])
def test_resample_reproduces_a_ricker_centred_on_a_node(dzw, dz):
    nw, center, nzp = 81, 40, 256
    w = _ricker(dzw * (np.arange(nw) - center)).astype(np.float32)
    out, _ = oracle.wavelet_resample(w, center, dzw, dz, nzp)
    z = _fft_order_z(dz, nzp).astype(np.float64)
    ind2 = np.floor(z.astype(np.float32) / np.float32(dzw))
    inside = (ind2 - 2 <= 0) & (ind2 + 3 >= 0)   # node 0 is one of the six
    assert inside.sum() >= 5
    np.testing.assert_allclose(out[inside], _ricker(z[inside]), atol=2e-7)
    # elsewhere the predictor is an interpolant of a well-sampled smooth pulse
    far = ~inside & (np.abs(z) < dzw * (nw // 2 - 4))
    assert np.abs(out[far] - _ricker(z[far])).max() < 2e-2


def test_resample_kriging_weights_vs_numpy_solve():
    rng = np.random.default_rng(2)
    nw, center, dzw, dz, nzp = 41, 20, 4.0, 1.7, 64
    w = rng.standard_normal(nw).astype(np.float32)
    out, _ = oracle.wavelet_resample(w, center, dzw, dz, nzp)
    z = _fft_order_z(dz, nzp)
    for k in range(nzp):
        i2 = int(np.floor(np.float32(z[k]) / np.float32(dzw)))
        ind = i2 + np.arange(6) - 2
        nodes = (np.float32(dzw) * ind.astype(np.float32)).astype(np.float32)
        C = _ricker((nodes[:, None] - nodes[None, :]).astype(np.float32).astype(np.float64))
        c = _ricker((nodes - np.float32(z[k])).astype(np.float32).astype(np.float64))
        ii = ind + center
        val = np.where((ii >= 0) & (ii < nw), w[np.clip(ii, 0, nw - 1)], 0).astype(np.float64)
        want = val @ np.linalg.solve(C, c)
        assert abs(out[k] - want) <= 1e-5 * max(1.0, abs(want))


def test_resample_is_linear():
    rng = np.random.default_rng(8)
    nw, center = 33, 16
    a = rng.standard_normal(nw).astype(np.float32)
    b = rng.standard_normal(nw).astype(np.float32)
    oa, _ = oracle.wavelet_resample(a, center, 4.0, 2.2, 50)
    ob, _ = oracle.wavelet_resample(b, center, 4.0, 2.2, 50)
    oab, _ = oracle.wavelet_resample((a + b).astype(np.float32), center, 4.0, 2.2, 50)
    np.testing.assert_allclose(oab, oa + ob, atol=5e-5)


def test_resample_reports_a_singular_kriging_system():
    # at 0.01 ms spacing the six Ricker correlations agree to ~1e-9: the
    # matrix is numerically singular and the reference's Cholesky throws
    w = np.ones(9, np.float32)
    assert oracle.wavelet_resample(w, 4, 0.01, 0.005, 16) is None


# --------------------------------------------------------------- energy
def _energy_inputs(nt=4, nzp=64, seed=0):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nt, nzp)).astype(np.float32)
    A = rng.standard_normal((nt, 3)).astype(np.float32)
    s = rng.standard_normal((3, 3))
    s0 = s @ s.T + np.eye(3)
    ev = (0.1 + rng.random(nt)).astype(np.float32)
    sn = (2.0 + 5 * rng.random(nt)).astype(np.float32)
    return w, A, s0, ev, sn


def test_energy_delta_correlation_is_the_sum_of_squared_differences():
    w, A, s0, ev, sn = _energy_inputs()
    nzp = w.shape[1]
    ct = np.zeros(nzp, np.float32)
    ct[0] = 1.0
    e = oracle.wavelet_energy(w, A, s0, ct, ev, sn)
    wd = np.roll(w, -1, axis=1).astype(np.float64) - w
    pv = np.einsum("la,ab,lb->l", A.astype(np.float64), s0, A.astype(np.float64))
    want = (wd ** 2).sum(axis=1) * pv
    np.testing.assert_allclose(e["model_var"], want, rtol=2e-5)
    np.testing.assert_allclose(e["signal_var"], ev + want, rtol=2e-5)
    np.testing.assert_allclose(e["theo_sn"], (ev + want) / ev, rtol=2e-5)
    assert np.all(e["gain"] == 1.0)


def test_energy_toeplitz_quadratic_form_and_matching():
    w, A, s0, ev, sn = _energy_inputs(nt=5, nzp=80, seed=4)
    nzp = w.shape[1]
    ct = np.exp(-np.arange(nzp) / 6.0).astype(np.float32)
    match = np.array([1, 0, 1, 1, 0], np.uint8)
    e = oracle.wavelet_energy(w, A, s0, ct, ev, sn, match)
    wd = np.roll(w, -1, axis=1).astype(np.float64) - w
    T = ct.astype(np.float64)[np.abs(np.arange(nzp)[:, None] - np.arange(nzp)[None, :])]
    pv = np.einsum("la,ab,lb->l", A.astype(np.float64), s0, A.astype(np.float64))
    mv = np.einsum("li,ij,lj->l", wd, T, wd) * pv
    g = np.sqrt(ev / mv * (sn - 1))
    np.testing.assert_allclose(e["gain"], np.where(match == 1, g, 1.0), rtol=2e-5)
    np.testing.assert_allclose(e["model_var"], np.where(match == 1, mv * g * g, mv), rtol=4e-5)
    # matched angles: theoretical S/N = empirical S/N
    np.testing.assert_allclose(e["theo_sn"][match == 1], sn[match == 1], rtol=2e-6)
    np.testing.assert_allclose(e["theo_sn"][match == 0], ((ev + mv) / ev)[match == 0], rtol=4e-5)
