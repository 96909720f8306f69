"""CPU checks of the C ABI: the library loads and exports every entry point
that include/molla_mpm.h declares; host-side validation paths need no GPU."""
import os
import re

import ctypes as C
import pytest

from molla_amd import mpm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "molla_mpm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", txt)))


def test_header_declarations_are_exported():
    names = _declared()
    assert len(names) >= 25
    L = C.CDLL(mpm.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert sorted(mpm.exported_symbols()) == _declared()


def test_abi_version_and_null_handles():
    L = mpm.lib()
    assert L.mpm_abi_version() == 1
    assert L.mpm_destroy(None) == 1
    assert L.mpm_step_n(None, C.c_float(1e-3), 1) == 1
    assert L.mpm_last_error(None) == b"null handle"
