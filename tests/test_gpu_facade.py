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


def test_python_simulation_relax_and_propagator(tdse, orc):
    """WavepacketSimulation.relax(n, dtau) (core/src/simulation.ixx:45-52) after
    advance(n) equals the oracle's Strang steps then relax steps, the clock
    counts only real time, and propagator() exposes the reference tables."""
    dims, lo, hi = (32, 16, 16), (-8.0, -6.0, -6.0), (8.0, 6.0, 6.0)
    grid = tdse.Grid3D(*[tdse.Grid1D(lo[a], hi[a], dims[a]) for a in range(3)])
    v = orc.harmonic_3d(dims, lo, hi, 0.8, (0.0, 0.0, 0.0))
    cfg = tdse.WavepacketSimulation.Config(grid, v, (1.0, 0.5, 0.0), (1.0, 1.0, 1.0),
                                           (0.2, 0.0, 0.0), 0.02)
    sim = tdse.WavepacketSimulation(cfg)
    sim.advance(5)
    sim.relax(40, 0.01)
    sim.relax(10, 0.01)
    got = sim.psi()
    psi0 = orc.gaussian_3d(dims, lo, hi, cfg.r0, cfg.sigma, cfg.k0).reshape(dims[::-1])
    want = orc.so3d_evolve(v, dims, lo, hi, 0.02, psi0, 5)
    want, _ = orc.it3d_relax(v, dims, lo, hi, 0.01, want, 50)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"case simulation relax: rel err {err:.3e}", flush=True)
    assert err < 1e-11
    assert sim.time() == 5 * 0.02
    half, kin = orc.so3d_tables(v, dims, lo, hi, 0.02)
    assert np.array_equal(sim.propagator().half_potential_phase().ravel(), half)
    assert np.array_equal(sim.propagator().kinetic_phase().ravel(), kin)
