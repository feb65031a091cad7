"""The C-ABI library builds, loads and exports every entry point declared in
include/cfd_abi.h (no compute calls: these run without a GPU)."""
import ctypes as C
import os
import subprocess

import pytest

from cfd_amd import _lib


def test_library_exports_every_declared_symbol():
    path = _lib.LIB_PATH
    assert os.path.isfile(path), "libcfd_amd.so not built"
    L = C.CDLL(path)
    names = _lib.declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(names) <= exported


def test_abi_version_and_enums():
    L = _lib.lib()
    assert L.cfd_abi_version() == 5
    assert _lib.FIELDS["CFD_F_U"] == 0 and _lib.FIELDS["CFD_NFIELDS"] == 34
    assert _lib.SCALARS["CFD_NSCALARS"] == 21
    assert C.sizeof(_lib.Consts) == 11 * 8


def test_library_has_gfx950_code_object():
    out = subprocess.run(["strings", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out
