"""SeismicParametersHolder::computeCircCorrT + makeCircCorrTPosDef
(src/seismicparametersholder.cpp:431-499) behind crava_hip_circ_corr_t
(host code, no GPU).

Pins:
* the reference's own steps on its own FFTW 2.1.2 float 1-D transforms
  (oracle/_ref, ref_circ_corr_t): within 2e-6 of c(0) = 1 (the library
  evaluates each transform in double and rounds once; FFTW's float
  transforms add a few ulps);
* closed forms: a correlation whose circular extension already has a
  nonnegative spectrum, with no bin removed, comes back as itself over its
  lag-0 value; the output's spectrum is real, nonnegative and zero at the
  removed low bins; c(0) = 1 to float rounding; the output is symmetric;
* the error paths: priorCorrT[0] == 0 (the reference asserts) and a
  spectrum left empty by the low cut (the reference exits)."""
import numpy as np
import pytest

import oracle
from crava_amd import _lib as L
from crava_amd.core import circ_corr_t


def spherical(n, r):
    t = np.arange(n) / r
    return np.where(t < 1, 1 - 1.5 * t + 0.5 * t ** 3, 0.0).astype(np.float32)


def gen_exp(n, r, p=1.5):
    return np.exp(-3 * (np.arange(n) / r) ** p).astype(np.float32)


CASES = [(spherical(40, 10.0), 2, 64), (spherical(256, 12.5), 5, 320), (gen_exp(70, 9.0), 0, 75),
         (gen_exp(200, 30.0, 1.0), 3, 243), (spherical(150, 20.0), 4, 160), (gen_exp(33, 4.0, 2.0), 1, 36)]


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("case", range(len(CASES)))
def test_against_reference_fftw(case):
    prior, cut, nzp = CASES[case]
    got = circ_corr_t(prior, cut, nzp)
    want, bad = oracle.circ_corr_t_ref(prior, cut, nzp)
    assert bad == 0
    assert np.abs(got - want).max() <= 2e-6, np.abs(got - want).max()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_closed_form_properties(case):
    prior, cut, nzp = CASES[case]
    got = circ_corr_t(prior, cut, nzp).astype(np.float64)
    assert abs(got[0] - 1.0) <= 2.0 ** -23
    k = np.arange(1, nzp)
    assert np.abs(got[k] - got[nzp - k]).max() <= 2e-6
    spec = np.fft.rfft(got)
    top = np.abs(spec).max()
    assert np.abs(spec.imag).max() <= 1e-5 * top
    assert spec.real.min() >= -1e-5 * top
    assert np.abs(spec[: cut + 1]).max() <= 1e-5 * top


def test_positive_spectrum_comes_back_unchanged():
    """A Gaussian correlation with a nugget at lag 0 over the whole circle
    has a spectrum of at least the nugget: |.| changes nothing, so with no
    bin removed (min_int_fq = -1) the output is the symmetric extension over
    its lag-0 value."""
    nzp = 96
    t = np.minimum(np.arange(nzp), nzp - np.arange(nzp))
    full = np.exp(-((t / 6.0) ** 2))
    full[0] += 0.1
    assert np.fft.rfft(full).real.min() > 0.09
    prior = full[: nzp // 2 + 1].astype(np.float32)
    got = circ_corr_t(prior, -1, nzp)
    assert np.abs(got - full / full[0]).max() <= 1e-6


def test_errors():
    with pytest.raises(L.CravaHipError):
        circ_corr_t(np.array([0.0, 0.5], np.float32), 0, 16)
    # a constant correlation has only a DC bin, which the cut removes
    with pytest.raises(L.CravaHipError):
        circ_corr_t(np.ones(9, np.float32), 0, 16)
