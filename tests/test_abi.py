"""The C-ABI library loads without a GPU and exports every entry point
that include/imago_hip.h declares; host-side argument checks."""
import os
import re
import subprocess

import pytest

from imago_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "imago_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(imago_\w+)\s*\(", src, re.M)))


def test_header_declares_the_expected_surface():
    syms = declared_symbols()
    assert set(L.EXPORTS) == set(syms), (set(L.EXPORTS) ^ set(syms))


def test_library_loads_and_exports_all_symbols():
    lib = L.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (imago_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_last_error_is_empty_string_initially():
    assert isinstance(L.load().imago_last_error(), bytes)


def test_context_on_missing_device_fails_loudly():
    # device index far out of range: must be an error, never a CPU fallback
    with pytest.raises(L.ImagoError):
        L.Context(987)
