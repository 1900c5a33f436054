"""The C-ABI library builds, loads and exports every entry point that
include/femhip.h declares.  CPU only: no compute calls."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "femhip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char|void|double)\s*\**\s*(femhip_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "femhip_solver_solve" in syms and "femhip_operator_compute" in syms
    assert len(syms) >= 21


def test_library_exports_every_declared_symbol():
    from fempar_amd import _native as N
    assert os.path.exists(N.LIB_PATH), "run fempar_amd.build.build() first"
    lib = ctypes.CDLL(N.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert sorted(N.EXPORTS) == sorted(s for s in declared_symbols() if s in N.EXPORTS)
    missing = [s for s in declared_symbols() if s not in N.EXPORTS]
    assert not missing, f"bindings lack {missing}"


def test_no_device_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from fempar_amd import _native as N
    lib = N.load()
    h = ctypes.c_void_p()
    rc = lib.femhip_ctx_create(0, ctypes.byref(h))
    assert rc == N.ERR_NO_DEVICE
    assert b"no HIP device" in lib.femhip_last_error()


def test_invalid_arguments_are_rejected_without_device():
    from fempar_amd import _native as N
    lib = N.load()
    assert lib.femhip_space_info(None, None, None, None, None, None) == N.ERR_INVALID
    assert b"null" in lib.femhip_last_error()
    assert lib.femhip_serve_submit(None, None, None) == N.ERR_INVALID


def test_serve_admission_estimate_bounds_the_stored_data():
    """The admission estimate of an elasticity request covers at least the
    stencil operator's FP64 and FP32 values (81 per row of vector Q1) and the
    Krylov vectors, and grows with n."""
    from fempar_amd import _native as N
    lib = N.load()
    est = [lib.femhip_serve_estimate_bytes(n, 20000) for n in (1, 8, 32, 96)]
    assert all(a < b for a, b in zip(est, est[1:]))
    for n, e in zip((8, 32, 96), est[1:]):
        ndof = 3 * n * (n + 1) ** 2
        assert e >= ndof * (81 * 12 + 8 * 8)
    assert est[-1] < 4e9   # the largest serving request (n = 96) stays well inside one GPU
