"""The oracle against the golden fixtures of tests/golden (outputs of the
reference's own FFTW 2.1.2 and lib_matr.c, written by
tests/golden/make_golden.py).  These pins hold where the reference sources
are absent."""
import os

import numpy as np
import pytest

import oracle
from tests.helpers import rel_l2

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_cases.npz"))


@pytest.mark.parametrize("i", range(4))
def test_chol_matches_reference_bitwise(i):
    a = G[f"chol{i}_in"].copy()
    st = oracle.chol_cpx(a)
    assert st == int(G[f"chol{i}_status"]) == 0
    assert np.array_equal(a.view(np.uint32), G[f"chol{i}_out"].view(np.uint32))


def test_chol_failure_status():
    a = G["chol_bad_in"].copy()
    assert oracle.chol_cpx(a) == int(G["chol_bad_status"]) != 0


@pytest.mark.parametrize("i", range(2))
def test_fft_matches_fftw(i):
    n = tuple(int(x) for x in G[f"fft{i}_shape"])
    got = oracle.fft_r2c(G[f"fft{i}_in"].copy(), *n, False)
    # double-precision oracle vs FFTW's float transform: float rounding only
    assert rel_l2(got, G[f"fft{i}_out"]) < 5e-7


def test_posterior_loop_matches_reference_bitwise():
    nxp, nyp, nzp, nt = (int(x) for x in G["post_dims"])
    tabs = {k[len("post_tab_"):]: G[k] for k in G.files if k.startswith("post_tab_")}
    mu = [g.copy() for g in G["post_in_mu"]]
    d = [g.copy() for g in G["post_in_d"]]
    cv = [g.copy() for g in G["post_in_cov"]]
    nf = oracle.posterior(nxp, nyp, nzp, G["post_A"], G["post_etc"], float(G["post_wnc"]), G["post_sigma0"], tabs,
                          mu, d, cv, G["post_in_ec"].copy())
    assert nf == int(G["post_nfail"])
    for got, ref in zip(mu + d + cv, list(G["post_out_mu"]) + list(G["post_out_d"]) + list(G["post_out_cov"])):
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
