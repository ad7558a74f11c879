"""The C++ facade (include/tdse/split_operator.hpp) from a reference-style call
site, on the GPU, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from test_abi import build_facade

pytestmark = pytest.mark.gpu


def test_cpp_facade_matches_oracle(orc, tmp_path):
    exe = build_facade(tmp_path)
    out = str(tmp_path / "psi.bin")
    r = subprocess.run([exe, out], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    dims, lo, hi = (32, 32, 16), (-8.0, -8.0, -4.0), (8.0, 8.0, 4.0)
    v = orc.harmonic_3d(dims, lo, hi, 0.5, (0.0, 0.0, 0.0))
    psi0 = orc.gaussian_3d(dims, lo, hi, (1.0, 0.0, 0.0), (1.0, 1.0, 1.0), (0.0, 0.5, 0.0))
    want = orc.so3d_evolve(v, dims, lo, hi, 0.02, psi0.reshape(16, 32, 32), 12).ravel()
    got = np.fromfile(out, dtype=np.complex128)
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-11
