"""The C-ABI library loads and exports every entry point include/lbm_hip.h
declares.  No device calls: runs without a GPU."""
import os
import re
import subprocess

import pytest

import lbm_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\s*\*?\s*(lbm_[a-z0-9_]+)\s*\(", txt, re.M))
    return names


def test_header_declares_entry_points():
    names = _declared()
    for must in ("lbm_create", "lbm_init", "lbm_collide", "lbm_stream", "lbm_macro", "lbm_step",
                 "lbm_download_macro", "lbm_export_links", "lbm_reduce_fp64", "lbm_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(lbm_amd.LIB_PATH):
        pytest.fail("liblbm_hip.so not built: run __graft_entry__.build()")
    out = subprocess.run(["nm", "-D", "--defined-only", lbm_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = _declared() - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_binding_covers_header_and_loads():
    lib = lbm_amd.load_library()
    assert set(lbm_amd.exported_symbols()) == _declared()
    # a pure host call: error strings
    assert lib.lbm_error_string(2).decode().startswith("density")


def test_library_is_gfx950_code_object():
    # the offload bundle of the library built here names its one target
    data = open(lbm_amd.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
