"""CPU tests of the drop-in boundary: the library builds and loads, exports every symbol
include/sep.h declares, rejects bad arguments without touching a GPU, and the host-side
mirror validates layouts the way the reference's API expects them.  No compute call is
made here (no GPU in this environment)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sep.h")


@pytest.fixture(scope="module")
def libpath(sep):
    if not os.path.exists(sep.LIB_PATH):
        pytest.fail(f"{sep.LIB_PATH} not built: run `python __graft_entry__.py` (build())")
    return sep.LIB_PATH


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sep_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_python_export_list(sep):
    assert declared_symbols() == sorted(sep.EXPORTS)


def test_every_declared_symbol_is_exported(libpath):
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_library_is_gfx950_only(libpath):
    blob = open(libpath, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    for other in (b"--gfx942", b"--gfx90a", b"--gfx1100"):
        assert other not in blob


def test_null_handle_is_rejected_without_gpu(sep):
    L = sep.lib()
    d = np.zeros(4)
    assert L.sep_qr_algorithm_f64(None, 4, d.ctypes.data, d.ctypes.data, None, 4, 128,
                                  1e-16, 400, None, None, None) == sep.SEP_ERR_ARG
    assert "null handle" in sep.last_error()
    assert L.sep_bulk_rmul_f64_dev(None, 4, 4, None, 4, None, 2) == sep.SEP_ERR_ARG
    assert L.sep_destroy(None) == sep.SEP_OK
    assert "gfx950" in sep.version()


def test_create_without_gpu_reports_error(sep):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = sep.lib().sep_create(C.byref(h), 0)
    assert rc < 0 and not h.value
    assert sep.last_error() != ""


def test_host_mirror_layout_checks(sep):
    H = sep.SymTridiagonal(np.ones(4), np.ones(3))
    with pytest.raises(ValueError):
        sep.SymTridiagonal(np.ones(4), np.ones(4))
    with pytest.raises(ValueError):
        sep._np_colmajor(np.ones((3, 3)))                  # C-ordered matrix
    assert sep._np_colmajor(np.asfortranarray(np.ones((3, 2)))) == (3, 2, 3)
    assert H.n == 4 and H.dense().shape == (4, 4)


def test_inputs_are_seeded_and_shaped(sep):
    d, e = sep.inputs.tridiag_linear(10)
    assert np.array_equal(d, np.arange(1.0, 11.0)) and e.shape == (9,)
    assert np.all((0 <= e) & (e < 1))
    g1, g2 = sep.inputs.givens_grid(7, 3, seed=5), sep.inputs.givens_grid(7, 3, seed=5)
    assert np.array_equal(g1, g2) and g1.shape == (3, 6, 2)
    assert np.allclose(g1[..., 0] ** 2 + g1[..., 1] ** 2, 1)
    ns = sep.inputs.batched_orders(4096)
    assert ns.min() >= 32 and ns.max() <= 512 and np.all(ns % 2 == 0)
    dd, ee, od, oq = sep.inputs.batched_problems(ns[:5])
    assert od[-1] + ns[4] == dd.shape[0] and oq[1] == ns[0] ** 2
    _, e2 = sep.inputs.tridiag_linear(20, e_low=0.1)
    assert e2.min() >= 0.1
