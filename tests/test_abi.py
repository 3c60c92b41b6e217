"""CPU checks of the C ABI boundary: the library loads and exports every
symbol include/qsim_hip.h declares.  No compute calls (no GPU here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qsim_hip.h")
LIB = os.path.join(ROOT, "qusim_amd", "libqsim_hip.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(qsim_hip_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "qsim_hip_split_steps" in names and "qsim_hip_fft_exec" in names
    assert len(names) > 60


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqsim_hip.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libqsim_hip.so not built")
def test_library_loads_and_binding_covers_header():
    from qusim_amd import _native
    L = _native.lib()
    for n in declared():
        assert hasattr(L, n), n
    assert set(declared()) <= set(_native.exported_symbols())
