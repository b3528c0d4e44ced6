"""The C ABI library: it loads without a GPU, exports every symbol that
include/mlegs_hip.h declares, and the Python binding covers exactly that set.
No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

from mlegs_amd import hip as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mlegs_hip.h")


def header_symbols():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\b(mlg_[a-z0-9_]+)\s*\(", src))


def test_library_built_and_loads():
    assert os.path.exists(H.library_path()), "run `make -C mlegs_amd` (or __graft_entry__.build())"
    lib = H._lib()
    assert lib.mlg_error_string(11) == b"SCALAR3_DEALIASING"
    assert lib.mlg_error_string(5) == b"SOLVER"


def test_every_header_symbol_exported():
    syms = header_symbols()
    assert len(syms) >= 50
    out = subprocess.check_output(["nm", "-D", "--defined-only", H.library_path()], text=True)
    exported = set(re.findall(r"\bT (mlg_[a-z0-9_]+)\b", out))
    assert syms <= exported, sorted(syms - exported)
    lib = ctypes.CDLL(H.library_path())
    for s in syms:
        getattr(lib, s)


def test_binding_covers_header():
    assert set(H.SIGNATURES) == header_symbols()


def test_product_does_not_reach_the_oracle():
    pkg = os.path.join(ROOT, "mlegs_amd")
    for dp, _, fns in os.walk(pkg):
        for fn in fns:
            if fn.endswith((".py", ".hip", ".cpp", ".h")):
                txt = open(os.path.join(dp, fn)).read()
                assert "_oracle" not in txt and "libmlegs_oracle" not in txt, fn
                assert not re.search(r"\bog_[a-z]", txt), fn


def test_params_struct_layout_matches_oracle():
    import _oracle as O
    assert [f[0] for f in H.Params._fields_] == [f[0] for f in O.OgParams._fields_]
    assert ctypes.sizeof(H.Params) == ctypes.sizeof(O.OgParams)


def test_fortran_binding_built_and_linked():
    drv = os.path.join(ROOT, "mlegs_amd", "fortran", "_build", "drive_bsnsq")
    assert os.path.exists(drv), "run `make -C mlegs_amd/fortran`"
    out = subprocess.check_output(["ldd", drv], text=True)
    line = [l for l in out.splitlines() if "libmlegs_hip.so" in l]
    assert line and "not found" not in line[0]
    assert os.path.realpath(line[0].split("=>")[1].split("(")[0].strip()) == os.path.realpath(H.library_path())
