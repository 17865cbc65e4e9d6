"""The C ABI driven from C++ alone (tests/cabi/harness.cpp, built by build()): a Model C
run in fused and per-intent mode with a Parquet observer, no Python in the loop.
The fused run equals the per-intent one and the oracle bit for bit, and the observer
wrote the static file and the snapshots of steps 1, 10 and 20."""
from __future__ import annotations

import json
import os
import pathlib
import subprocess
import sys

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

HARNESS = pathlib.Path(__file__).resolve().parent / "cabi" / "harness"


def test_cxx_caller_matches_the_oracle(orc, tmp_path):
    print("[case] C ABI from C++: 256x192 Model C, 25 steps", flush=True)
    sys.stdout.flush()
    assert HARNESS.exists(), f"{HARNESS} is missing: run __graft_entry__.build()"
    res = subprocess.run([str(HARNESS), str(tmp_path), "25"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["steps"] == 25 and out["fused_equals_per_intent"] == 1
    assert out["written"] == 4 and out["dropped"] == 0 and out["failed"] == 0
    assert sorted(os.listdir(tmp_path / "parquet")) == ["snap_0000001.parquet", "snap_0000010.parquet",
                                                       "snap_0000020.parquet", "static_0000000.parquet"]
    cells, block = [256, 192], [64, 64]
    og = orc.grid(cells, block, [0.0, 0.0], [0.4, 0.4])
    phi = np.zeros(orc.field_len(og, 1))
    u = np.zeros_like(phi)
    orc.initialize_grains(og, phi, u, np.array([[20.0, 30.0, 2.0], [70.0, 45.0, 2.0]]), 0.55)
    t = orc.modelc_run(og, orc.params(0.05, 3.19, 2.0), phi, u, 25, 0.008, seed=42)
    assert out["time"] == t
    got_phi = np.fromfile(tmp_path / "phi.bin", dtype="<f8")
    got_u = np.fromfile(tmp_path / "u.bin", dtype="<f8")
    assert np.array_equal(H.field_to_global(got_phi, cells, block, 1), H.field_to_global(phi, cells, block, 1))
    assert np.array_equal(H.field_to_global(got_u, cells, block, 1), H.field_to_global(u, cells, block, 1))
