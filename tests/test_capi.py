"""The C-ABI library loads and exports every symbol include/picamd.h declares (CPU)."""
import os
import re
import subprocess

import pytest

import pic_amd

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "picamd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(pic_\w+)\(", text, re.M)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(pic_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(pic_amd.LIB_PATH), "libpicamd.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", pic_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (pic_\w+)$", out, re.M))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_library_loads_and_reports_version():
    L = pic_amd.lib()
    want = int(re.search(r"#define PICAMD_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert L.pic_abi_version() == want
    for s in pic_amd.EXPORTS:
        assert hasattr(L, s)


def test_device_count_callable_without_gpu():
    assert pic_amd.device_count() >= 0
