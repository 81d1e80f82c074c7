"""CPU-side checks of the product library: it is built, loads, and exports every entry point the
public header declares (no GPU call is made here); the host-side Python mirror imports."""
import ctypes
import os

import pytest

from tessera_amd import lib as L


def test_library_exports_every_declared_symbol():
    assert os.path.exists(L.LIB_PATH), "build with `python -m tessera_amd.build`"
    so = ctypes.CDLL(L.LIB_PATH)
    declared = L.declared_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(so, s)]
    assert not missing, missing


def test_bindings_cover_the_header():
    declared = set(L.declared_symbols())
    assert declared == set(L._SIGS), (declared - set(L._SIGS), set(L._SIGS) - declared)


def test_version_and_error_without_gpu():
    lib = L.lib()
    assert b"gfx950" in lib.tsr_version()
    assert isinstance(lib.tsr_last_error(), bytes)


def test_oracle_has_no_product_dependency():
    """The product never loads the oracle: no file under tessera_amd/ mentions it."""
    root = os.path.dirname(L.HERE)
    for dirpath, _, files in os.walk(os.path.join(root, "tessera_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "liboracle" not in text and "from oracle" not in text and "import oracle" not in text, f
