"""C ABI surface, host-side entry points and argument checking (no GPU)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import oracle
from crava_amd import _lib as L
from crava_amd import synth
from tests.helpers import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(crava_hip_\w+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 15
    lib = L.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_declarations():
    assert set(declared_functions()) <= set(L.SIGNATURES)


def test_version_and_error_channel():
    lib = L.lib()
    assert lib.crava_hip_version() >= 100
    assert lib.crava_hip_ctx_create(None, None) == L.E_INVALID
    assert b"NULL" in lib.crava_hip_last_error()
    assert lib.crava_hip_fft_r2c(None, None, 8, 8, 8, 0, 0) == L.E_INVALID
    assert lib.crava_hip_post_run(None, 0, None, None, None, None, None, None) == L.E_INVALID
    assert lib.crava_hip_facies(None, 1, 1, 1, 1, 1, 1, None, None, None, 1, None, 1, 1, 1, None, None, None,
                                0.0, None, None, None) == L.E_INVALID


@pytest.mark.parametrize("name", ["small", "plain_cpu", "precision"])
def test_perk_tables_host_code_vs_oracle(name):
    """The library's host-side per-k preparation against the oracle's."""
    from crava_amd.core import perk_tables
    prob = synth.make_problem(synth.config(name), grids=False)
    cfg = prob.cfg
    args = (prob.ntheta, cfg.nz, cfg.nzp, prob.lz, cfg.low_cut, cfg.high_cut, 1.0, prob.wavelets, prob.wnorm)
    got = perk_tables(*args)
    ref = oracle.perk_tables(*args)
    assert np.array_equal(got["invert"], ref["invert"])
    for key in ("kWD", "e1", "e2", "e3"):
        assert rel_l2(got[key], ref[key]) < 1e-6, key


def test_perk_tables_rejects_bad_args():
    lib = L.lib()
    z = np.zeros(64, np.float32)
    rc = lib.crava_hip_perk_tables(9, 8, 8, 32.0, 5.0, 55.0, 1.0, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                                   z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data)
    assert rc == L.E_INVALID


def test_product_never_imports_the_oracle():
    for path in glob.glob(os.path.join(ROOT, "crava_amd", "**", "*"), recursive=True):
        if os.path.isfile(path) and path.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
            text = open(path).read()
            assert "import oracle" not in text and "from oracle" not in text, path
            assert "liboracle" not in text and "libcrava_ref" not in text, path


def test_host_mirror_links_only_the_c_abi():
    """The C++ host mirror (FFTGrid / SeismicParametersHolder / Crava) calls
    the GPU only through include/crava_hip.h: every undefined symbol it needs
    from libcrava_hip.so is a declared crava_hip_* entry point, and it needs
    nothing from the HIP runtime directly."""
    import subprocess
    so = os.path.join(ROOT, "crava_amd", "libcrava_host.so")
    assert os.path.exists(so), "built by crava_amd.build"
    out = subprocess.run(["nm", "-D", "--undefined-only", so], capture_output=True, text=True, check=True).stdout
    undef = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    ours = {u for u in undef if u.startswith("crava_hip_")}
    assert ours and ours <= set(declared_functions())
    assert not any(u.startswith("hip") for u in undef)
    C.CDLL(so)   # loads with its rpath to libcrava_hip.so


@pytest.mark.parametrize("cname,pyname", [("crava_hip_post_params", "PostParams"),
                                          ("crava_hip_batch_volume", "BatchVolume"),
                                          ("crava_hip_synt_params", "SyntParams"),
                                          ("crava_hip_swdiv_params", "SwdivParams"),
                                          ("crava_hip_local_noise", "LocalNoiseParams")])
def test_struct_layouts_match_the_header(cname, pyname, tmp_path):
    """Every ctypes mirror of a header struct has the header's size and field
    offsets (gcc on the header)."""
    import subprocess
    py = getattr(L, pyname)
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "crava_hip.h"', "int main(void) {",
             f'  printf("%zu\\n", sizeof({cname}));']
    lines += [f'  printf("%zu\\n", offsetof({cname}, {f[0]}));' for f in py._fields_]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [C.sizeof(py)] + [getattr(py, f[0]).offset for f in py._fields_]


def test_preparation_entry_points_reject_null_arguments():
    lib = L.lib()
    assert lib.crava_hip_divide_by_scale_wavelet(None, None, None) == L.E_INVALID
    assert lib.crava_hip_interpolate_seismic(None, 9, 8, 20, 12, 11, 24, 0.1, None, None, None) == L.E_INVALID


@pytest.mark.parametrize("n", [(10, 6, 8), (9, 7, 5), (64, 50, 40)])
def test_reference_noise_stream_matches_the_oracle(n):
    """crava_hip_complex_noise_reference reproduces the reference's RandomGen
    stream in FFTGrid::fillInComplexNoise's order bit for bit (the oracle's
    restatement, pinned by the LCG closed form and the normal moments), and
    hands back the advanced generator state; two calls continue the stream."""
    nxp, nyp, nzp = n
    cn = nxp // 2 + 1
    lib = L.lib()
    for seed in (23665, 99, 4000000001):
        s = C.c_uint32(seed)
        out = np.empty((nzp, nyp, 2 * cn), np.float32)
        assert lib.crava_hip_complex_noise_reference(nxp, nyp, nzp, C.byref(s), out.ctypes.data) == L.OK
        want, s_after = oracle.fill_complex_noise(nxp, nyp, nzp, seed)
        assert np.array_equal(out, want) and s.value == s_after
        out2 = np.empty_like(out)
        assert lib.crava_hip_complex_noise_reference(nxp, nyp, nzp, C.byref(s), out2.ctypes.data) == L.OK
        want2, _ = oracle.fill_complex_noise(nxp, nyp, nzp, s_after)
        assert np.array_equal(out2, want2)
    assert lib.crava_hip_complex_noise_reference(0, 1, 1, None, None) == L.E_INVALID
