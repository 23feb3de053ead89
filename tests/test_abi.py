"""CPU-side checks of the C ABI: the built library loads and exports every
entry point that include/dwf_gpu.h declares (no compute calls)."""
import os
import re
import subprocess

import pytest

from cps_veloc_amd import dwf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header=None):
    names = set()
    for h in ([header] if header else ["dwf_gpu.h", "dwf_synth.h"]):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(dwf(?:gpu|synth)_\w+)\s*\(", text))
    return sorted(names)


def test_header_matches_binding_list():
    assert declared("dwf_gpu.h") == sorted(dwf.EXPORTS)
    assert declared("dwf_synth.h") == sorted(dwf.SYNTH_EXPORTS)


@pytest.mark.skipif(not os.path.exists(dwf.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", dwf.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT\s+(dwf(?:gpu|synth)_\w+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing
    lib = dwf.lib()
    for s in declared():
        assert getattr(lib, s) is not None


@pytest.mark.skipif(not os.path.exists(dwf.LIB_PATH), reason="library not built")
def test_library_is_gfx950():
    """The embedded code object targets gfx950 (our own build's fat binary)."""
    data = open(dwf.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
