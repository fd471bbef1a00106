"""The C ABI library loads and exports every symbol the headers declare.

CPU only: no compute call is made (there is no GPU here)."""
import ctypes
import os
import re

import pytest

import flip3d_amd as fa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_symbols():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(flip_(?:gpu|unet)_\w+)\s*\(", text))
    return names


def test_headers_declare_the_abi():
    names = declared_symbols()
    assert "flip_gpu_step" in names and "flip_gpu_solve" in names and "flip_gpu_map_p2g" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(fa.LIB_PATH):
        pytest.fail(f"{fa.LIB_PATH} not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(fa.LIB_PATH)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(fa.EXPORTS) <= declared_symbols()


def test_default_params_match_reference_defaults():
    """flip_gpu_default_params needs no device: the reference's #defines."""
    p = fa.default_params(32)
    assert (p.n, p.dt, p.alpha, p.gravity, p.density, p.subcell) == (32, 0.6e-2, 0.95, 9.8, 0.5, 1)
    assert p.tol_policy == 0 and p.precond == fa.PRECOND_REFERENCE and p.rand_seed == 1


def test_errors_come_back_as_codes():
    """Without a device the create call fails with FLIP_E_HIP and a message
    (on a GPU host it succeeds); nothing throws across the ABI."""
    lib = fa.lib()
    p = fa.default_params(1)  # n out of range -> FLIP_E_ARG before any device call
    h = ctypes.c_void_p()
    rc = lib.flip_gpu_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1
    assert b"n out of range" in lib.flip_gpu_last_error()
