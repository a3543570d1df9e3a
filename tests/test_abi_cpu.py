"""CPU-side checks of the C ABI: the library loads and exports every entry point
the public header declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "admp_hip.h")
LIB = os.path.join(ROOT, "dmff_amd", "libadmp_hip.so")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(admp_[a-z_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("library not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def test_header_declares_expected_entry_points():
    d = declared()
    for name in ["admp_ctx_create", "admp_set_topology", "admp_set_pairs", "admp_build_neighbors",
                 "admp_compute", "admp_disp_compute", "admp_slater_compute", "admp_ctx_destroy"]:
        assert name in d


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing
    for s in declared():
        assert hasattr(lib, s)


def test_python_binding_matches_header():
    from dmff_amd import _lib
    assert sorted(_lib.SYMBOLS) == declared()
    # struct layouts: the ctypes mirrors have the field count of the C structs
    txt = open(HDR).read()
    body = re.search(r"typedef struct \{(.*?)\} admp_config;", txt, re.S).group(1)
    nfields = sum(len(line.split(";")) - 1 for line in body.splitlines() if ";" in line and not line.strip().startswith("/*"))
    nf = 0
    for line in body.splitlines():
        code = line.split("/*")[0]
        if ";" in code:
            nf += len([x for x in code.split(";")[0].split(",")])
    assert nf == len(_lib.Config._fields_)
    assert nfields >= 1


def test_version_string_without_gpu(lib):
    lib.admp_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.admp_version()
