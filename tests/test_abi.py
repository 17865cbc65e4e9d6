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


# ctypes mirror of each struct the header declares.
STRUCTS = {
    "lgd_cartesian_desc": "CartesianDesc",
    "lgd_face_bc": "FaceBcC",
    "lgd_modelc_params": "ModelCParamsC",
    "lgd_regrid_policy": "RegridPolicyC",
    "lgd_realization": "RealizationC",
    "lgd_tuning_result": "TuningResultC",
    "lgd_amr_tuning_result": "AmrTuningResultC",
    "lgd_observed_field": "ObservedFieldC",
    "lgd_driver": "DriverC",
}


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every struct of the header has a ctypes mirror with the same size and field
    offsets as the C compiler lays it out."""
    from legendre_amd import _abi

    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        st = getattr(_abi, pyname)
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in st._fields_:
            # a Python keyword in C (lambda) is mirrored with a trailing underscore
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field.rstrip("_")}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    res = subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src),
                          "-o", str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    c = {tuple(line.split()[:2]): int(line.split()[2]) for line in got if line}
    for cname, pyname in STRUCTS.items():
        st = getattr(_abi, pyname)
        assert c[(cname, "size")] == __import__("ctypes").sizeof(st), cname
        for field, _ in st._fields_:
            assert c[(cname, field)] == getattr(st, field).offset, (cname, field)
