"""The C-ABI library loads and exports every entry point include/cpet_hip.h
declares (no compute calls; this runs without a GPU)."""
import ctypes
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "cpet_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|char\s*\*|const\s+char\s*\*)\s*\*?\s*(\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_declares_reference_names():
    names = declared_functions()
    for n in ["compute_looped_field", "calc_esp_base", "calc_field", "calc_field_base",
              "thread_operation", "cpet_field_grid", "cpet_field_points", "cpet_topo",
              "cpet_ctx_create", "cpet_last_error"]:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    from pycpet_amd import _native as nat

    if not nat.LIB_PATH.exists():
        pytest.fail("libcpet_hip.so is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", str(nat.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    lib = nat.load()
    for n in declared_functions():
        assert hasattr(lib, n)
    assert set(nat.SIGNATURES) | {"cpet_last_error"} >= set(declared_functions())


def test_gpu_code_object_targets_gfx950():
    from pycpet_amd import _native as nat

    # the offload bundle of our own build names its target triple
    assert b"amdgcn-amd-amdhsa--gfx950" in nat.LIB_PATH.read_bytes()
