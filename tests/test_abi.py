"""The C-ABI library loads and exports every symbol include/afqmc_hip.h declares
(no GPU needed: nothing that reaches a device is called)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "afqmc_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void|size_t|const char\*)\s+(afqmc_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "afqmc_propagate" in syms and "afqmc_local_energy" in syms and len(syms) >= 15


def test_library_exports_every_declared_symbol():
    from ipie_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libafqmc_hip.so is not built (run python -m ipie_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    # the Python binding declares exactly the header's entry points
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "ipie_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle\b", src, flags=re.M), f


def test_mixed_context_refuses_a_taylor_order_beyond_its_term_buffers():
    """The split-BF16 Taylor series keeps every term's planes until the last
    term (at most 24 terms): afqmc_create refuses a larger exp_nmax in mixed
    precision with AFQMC_ERR_SHAPE, before it touches a device."""
    from ipie_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    lib.afqmc_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(_lib.Dims)]
    lib.afqmc_create.restype = ctypes.c_int
    ctx = ctypes.c_void_p()
    dims = _lib.Dims(24, 72, 5, 5, 10, 25, 0, _lib.PRECISIONS["mixed"])
    assert lib.afqmc_create(ctypes.byref(ctx), 0, ctypes.byref(dims)) == 2
    assert ctx.value is None
