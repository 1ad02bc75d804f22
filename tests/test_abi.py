"""CPU-side checks of the C-ABI library: it exists, loads, and exports every
symbol include/fcp.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIB = os.path.join(ROOT, "freecappuccino-dev_amd", "lib", "libfcp.so")
HDR = os.path.join(ROOT, "include", "fcp.h")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "freecappuccino-dev_amd", "csrc")],
                       check=True)
    return ctypes.CDLL(LIB)


def declared_symbols():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fcp_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("fcp_init", "fcp_mesh_upload", "fcp_geometry", "fcp_csr_build", "fcp_laplacian",
              "fcp_grad_gauss", "fcp_csrsolve", "fcp_iccg", "fcp_dpcg", "fcp_bicgstab"):
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_signatures_cover_header():
    import fcp._lib as L
    assert set(declared_symbols()) <= set(L.SIGNATURES), set(declared_symbols()) - set(L.SIGNATURES)


def test_library_is_gfx950_code_object(lib):
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "--notes", LIB], capture_output=True,
                         text=True)
    # the fat binary embeds gfx950 code objects
    data = open(LIB, "rb").read()
    assert b"gfx950" in data
