"""crava_hip_reflection_matrix (ModelAVODynamic::setupDefaultReflectionMatrix,
src/modelavodynamic.cpp:876-928) against what fixes it independently:

* PP is the linearised Zoeppritz equation in the Aki-Richards form
  R(t) = (1 + tan^2 t)/2 dVp/Vp - 4 g^2 sin^2 t dVs/Vs + (1 - 4 g^2 sin^2 t)/2 drho/rho,
  g = Vs/Vp: at normal incidence [1/2, 0, 1/2]; the Vp term is 1/(2 cos^2 t);
  the rho term equals 1/2 plus half the Vs term;
* PS: no Vp term, zero at normal incidence, and the Vs term equals the
  Aki-Richards converted-wave coefficient
  4 sin t / (2 cos j) (g^2 sin^2 t - g cos t cos j), sin j = g sin t.  The
  density term follows the reference as written (its cos t cos j term has
  the opposite sign to Aki-Richards' form); its g -> 0 limit -sin(t)/2 and
  small-angle slope (2g - 1)/2 hold either way.
"""
import math

import numpy as np
import pytest

from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.core import reflection_matrix

TH = np.radians(np.array([0.0, 4.0, 12.5, 25.0, 40.0, 55.0], dtype=np.float64)).astype(np.float32)


@pytest.mark.parametrize("g", [0.35, 0.5, 0.6])
def test_pp_closed_forms(g):
    A = reflection_matrix(TH, g).astype(np.float64)
    assert np.array_equal(A[0], [0.5, 0.0, 0.5])
    t = TH.astype(np.float64)
    assert np.allclose(A[:, 0] * 2 * np.cos(t) ** 2, 1.0, rtol=1e-6)
    assert np.allclose(A[:, 1], -4 * g * g * np.sin(t) ** 2, rtol=1e-6, atol=1e-9)
    assert np.allclose(A[:, 2] - A[:, 1] / 2, 0.5, rtol=0, atol=2e-7)


def test_pp_matches_the_workload_generator():
    deg = [4.0, 13.0, 22.0, 31.0, 40.0]
    A = reflection_matrix(np.radians(deg), 0.5)
    assert np.allclose(A, synth.reflection_matrix(deg, 0.5), rtol=2e-7, atol=1e-8)


@pytest.mark.parametrize("g", [0.35, 0.5, 0.6])
def test_ps_closed_forms(g):
    A = reflection_matrix(TH, g, seismic_type=[1] * len(TH)).astype(np.float64)
    t = TH.astype(np.float64)
    assert np.array_equal(A[:, 0], np.zeros(len(t)))
    assert np.array_equal(A[0], [0.0, 0.0, 0.0])
    cj = np.sqrt(1 - (g * np.sin(t)) ** 2)
    ar_vs = 4 * np.sin(t) / (2 * cj) * (g * g * np.sin(t) ** 2 - g * np.cos(t) * cj)
    assert np.allclose(A[:, 1], ar_vs, rtol=1e-6, atol=1e-9)
    # small-angle slope of the density term
    small = reflection_matrix(np.array([1e-3], np.float32), g, [1])[0, 2]
    assert abs(small / 1e-3 - (2 * g - 1) / 2) < 1e-4


def test_ps_density_term_at_vanishing_vs():
    A = reflection_matrix(TH, 1e-9, seismic_type=[1] * len(TH)).astype(np.float64)
    assert np.allclose(A[:, 2], -0.5 * np.sin(TH.astype(np.float64)), rtol=1e-6, atol=1e-9)


def test_mixed_types_and_bad_arguments():
    A = reflection_matrix(TH[:3], 0.5, seismic_type=[0, 1, 0])
    assert A[1, 0] == 0.0 and A[2, 0] != 0.0
    with pytest.raises(L.CravaHipError):
        reflection_matrix(TH[:2], 0.5, seismic_type=[0, 2])
    assert L.lib().crava_hip_reflection_matrix(0, None, None, 0.5, None) == L.E_INVALID
