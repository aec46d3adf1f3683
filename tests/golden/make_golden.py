"""Writes the golden fixtures of this directory from the reference's own C
sources: FFTW 2.1.2 (float) and libs/lib/lib_matr.c, compiled from
/root/reference by oracle/Makefile into oracle/_ref (this script needs that
library, so it runs where the reference sources are).  The fixtures are
data only (inputs and the reference's outputs), so the oracle can be
pinned where the reference sources are absent.

    python tests/golden/make_golden.py [reference|forward]   (default: both)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from crava_amd import synth  # noqa: E402
from tests.helpers import fourier_inputs, seeded_real_grid  # noqa: E402


def herm_pd(n, rng, cond):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return (Q @ np.diag(np.geomspace(1.0, cond, n)) @ Q.conj().T).astype(np.complex64)


def main():
    assert oracle.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    rng = np.random.default_rng(2024)
    out = {}
    # lib_matrCholCpx (libs/lib/lib_matr.c:78-167): in place, status
    for i, (n, cond) in enumerate([(1, 1.0), (3, 10.0), (6, 1e3), (8, 1e5)]):
        a = herm_pd(n, rng, cond)
        b = a.copy()
        st = oracle.chol_cpx(b, "ref")
        out[f"chol{i}_in"], out[f"chol{i}_out"], out[f"chol{i}_status"] = a, b, np.int32(st)
    bad = herm_pd(4, rng, 10.0)
    bad[2, 2] = -1.0
    b = bad.copy()
    out["chol_bad_in"], out["chol_bad_status"] = bad, np.int32(oracle.chol_cpx(b, "ref"))
    # rfftwnd 3-D r2c (FFTW 2.1.2 float) of seeded grids, unitary scaling
    for i, n in enumerate([(15, 12, 24), (9, 8, 10)]):
        g = seeded_real_grid(*n, seed=40 + i)
        out[f"fft{i}_shape"] = np.array(n, np.int32)
        out[f"fft{i}_in"] = g.copy()
        out[f"fft{i}_out"] = oracle.fft_r2c(g, *n, False, "ref")
    # the per-point posterior loop of Crava::computePostMeanResidAndFFTCov on
    # the reference's lib_matr, on FFTW-transformed inputs of a small problem
    prob = synth.make_problem(synth.config("small"))
    c = prob.cfg
    tabs = oracle.perk_tables(prob.ntheta, c.nz, c.nzp, prob.lz, c.low_cut, c.high_cut, 1.0, prob.wavelets,
                              prob.wnorm)
    mu, d, cv, ec = fourier_inputs(prob, "ref")
    out["post_dims"] = np.array([c.nxp, c.nyp, c.nzp, prob.ntheta], np.int32)
    out["post_A"], out["post_etc"], out["post_sigma0"] = prob.A, prob.err_theta_cov, prob.sigma0
    out["post_wnc"] = np.float32(c.wnc)
    for k, v in tabs.items():
        out[f"post_tab_{k}"] = v
    for name, lst in (("mu", mu), ("d", d), ("cov", cv)):
        out[f"post_in_{name}"] = np.stack(lst)
    out["post_in_ec"] = ec.copy()
    nf = oracle.posterior(c.nxp, c.nyp, c.nzp, prob.A, prob.err_theta_cov, c.wnc, prob.sigma0, tabs, mu, d, cv, ec,
                          "ref")
    for name, lst in (("mu", mu), ("d", d), ("cov", cv)):
        out[f"post_out_{name}"] = np.stack(lst)
    out["post_nfail"] = np.int64(nf)
    np.savez_compressed(os.path.join(HERE, "reference_cases.npz"), **out)
    print("wrote", os.path.join(HERE, "reference_cases.npz"))


def forward():
    """Crava::computeSyntSeismic's trace loop on the reference's FFTW 2.1.2
    (oracle/ref_driver.c): a small odd-sized volume, three angles, local
    shifts, gains and stretches on some angles."""
    assert oracle.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    rng = np.random.default_rng(77)
    nx, ny, nz, nxp, nyp, nzp = 7, 6, 21, 9, 8, 40
    rnxp = 2 * (nxp // 2 + 1)
    out = {"dims": np.array([nx, ny, nz, nxp, nyp, nzp], np.int32), "dz": np.float32(4.0)}
    for name in ("alpha", "beta", "rho"):
        out[name] = rng.standard_normal((nzp, nyp, rnxp)).astype(np.float32)
    out["A"] = rng.standard_normal((3, 3)).astype(np.float32)
    t = np.arange(nzp)
    t = np.minimum(t, nzp - t).astype(np.float64)
    out["wavelets"] = np.stack([np.exp(-(t / (2.0 + l)) ** 2) * np.cos(0.7 * t * (l + 1))
                                for l in range(3)]).astype(np.float32)
    out["shift0"] = rng.uniform(-10, 10, (ny, nx)).astype(np.float32)
    out["gain1"] = rng.uniform(0.5, 1.5, (ny, nx)).astype(np.float32)
    out["stretch0"] = rng.uniform(0.7, 1.3, (ny, nx)).astype(np.float32)
    out["stretch2"] = rng.uniform(0.9, 1.1, (ny, nx)).astype(np.float32)
    tabs = {name: [out.get(f"{name}{l}") for l in range(3)] for name in ("shift", "gain", "stretch")}
    res = oracle.synthetic_seismic(nx, ny, nz, nxp, nyp, nzp, out["A"], out["wavelets"], 4.0, out["alpha"],
                                   out["beta"], out["rho"], which="ref", **tabs)
    for l, g in enumerate(res):
        out[f"out{l}"] = g
    np.savez_compressed(os.path.join(HERE, "forward_cases.npz"), **out)
    print("wrote", os.path.join(HERE, "forward_cases.npz"))


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which in ("all", "reference"):
        main()
    if which in ("all", "forward"):
        forward()
