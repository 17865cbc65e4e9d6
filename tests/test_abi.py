"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
entry point include/legendre_gpu.h declares (and the ctypes table binds exactly
that set), and the product package never reaches into the oracle."""
from __future__ import annotations

import pathlib
import re
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "legendre_gpu.h"


def declared_functions() -> set:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(lgd_[a-z0-9_]+)\s*\(", text))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("lgd_ctx_create", "lgd_buf_alloc", "lgd_fill_ghosts", "lgd_modelc_step",
                     "lgd_sim_step", "lgd_apply_wiener"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from legendre_amd import _abi

    lib = _abi.lib()  # loads without a GPU; no compute call is made
    out = subprocess.run(["nm", "-D", "--defined-only", str(_abi.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (lgd_[a-z0-9_]+)$", out, flags=re.M))
    missing = declared_functions() - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    for name in declared_functions():
        assert getattr(lib, name) is not None
    assert lib.lgd_abi_version() == 1


def test_ctypes_table_matches_the_header():
    from legendre_amd import _abi

    assert set(_abi.SIGNATURES) == declared_functions()


def test_product_package_never_imports_the_oracle():
    pkg = ROOT / "legendre_amd"
    for path in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.h")):
        text = path.read_text()
        assert not re.search(r"^\s*(import|from)\s+oracle\b", text, flags=re.M), path
        assert "legendre_oracle" not in text, path


def test_library_links_no_oracle_code():
    from legendre_amd import _abi

    out = subprocess.run(["nm", "-D", str(_abi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    assert "orc_" not in out
