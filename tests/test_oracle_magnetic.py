"""The oracle's restatement of the three-shear rotation and the magnetic
split-operator step (core/src/rotation.ixx, core/src/magnetic.ixx), pinned by
the reference's tests/rotation_test.cpp and tests/magnetic_test.cpp closed
forms and, where the reference core is built, bit for bit."""
import math
import os

import numpy as np
import pytest

needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                    "libsesref.so")),
    reason="reference core not built here (oracle/Makefile ref needs the reference sources)")


def cube(half, n):
    return (n, n, n), (-half,) * 3, (half,) * 3


def blob(orc, d, lo, hi, x0, y0, sigma):
    x = lo[0] + np.arange(d[0]) * (hi[0] - lo[0]) / d[0]
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    f = np.exp(-((x3 - x0) ** 2 + (y3 - y0) ** 2 + z3 ** 2) / (2 * sigma * sigma)).astype(np.complex128)
    dv = ((hi[0] - lo[0]) / d[0]) ** 3
    return f / math.sqrt(orc.norm_sq(f, dv)), dv


def test_rotation_conserves_norm_and_rotates_the_mean(orc):
    d, lo, hi = cube(16.0, 64)
    f, dv = blob(orc, d, lo, hi, 4.0, 0.0, 1.5)
    n0 = orc.norm_sq(f, dv)
    r = orc.rotate_axis(f, d, lo, hi, 2, 0.6)
    assert abs(orc.norm_sq(r, dv) - n0) < 1e-11
    m = orc.mean_position_3d(r, d, lo, hi)
    assert abs(m[0] - 4 * math.cos(0.6)) < 2e-2 and abs(m[1] - 4 * math.sin(0.6)) < 2e-2
    assert abs(m[2]) < 1e-6
    fy, _ = blob(orc, d, lo, hi, 0.0, 4.0, 1.5)
    rx = orc.rotate_axis(fy, d, lo, hi, 0, 0.6)
    m = orc.mean_position_3d(rx, d, lo, hi)
    assert abs(m[0]) < 1e-6
    assert abs(m[1] - 4 * math.cos(0.6)) < 2e-2 and abs(m[2] - 4 * math.sin(0.6)) < 2e-2


def test_magnetic_effective_potential_and_zero_field(orc):
    d, lo, hi = cube(12.0, 32)
    v = orc.soft_coulomb_3d(d, lo, hi, 1.0, 1.0, (0, 0, 0))
    x = -12.0 + np.arange(32) * 0.75
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    for axis, perp2 in ((2, x3 ** 2 + y3 ** 2), (0, y3 ** 2 + z3 ** 2)):
        veff = orc.magnetic_veff(v, d, lo, hi, 0.4, axis)
        assert np.abs(veff - (v + 0.4 * 0.4 / 8.0 * perp2.ravel())).max() < 1e-12
    psi0 = orc.gaussian_3d(d, lo, hi, (2.0, 1.0, 0.0), (1.5, 1.5, 1.5), (0.0, 0.3, 0.0))
    mag = orc.magnetic_evolve(v, d, lo, hi, 0.01, 0.0, 2, psi0, 30)
    plain = orc.so3d_evolve(v, d, lo, hi, 0.01, psi0, 30)
    assert np.abs(mag - plain).max() < 1e-10


def test_magnetic_factors_into_core_plus_larmor_rotation(orc):
    d, lo, hi = cube(12.0, 32)
    v = orc.soft_coulomb_3d(d, lo, hi, 1.0, 1.0, (0, 0, 0))
    dt, b, n = 0.01, 0.5, 40
    for axis, r0, k0 in ((2, (2.5, 0.0, 0.5), (0.0, 0.4, 0.0)), (0, (0.5, 2.5, 0.0), (0.4, 0, 0))):
        psi0 = orc.gaussian_3d(d, lo, hi, r0, (1.4, 1.4, 1.4), k0)
        mag = orc.magnetic_evolve(v, d, lo, hi, dt, b, axis, psi0, n)
        veff = orc.magnetic_veff(v, d, lo, hi, b, axis)
        ref = orc.so3d_evolve(veff, d, lo, hi, dt, psi0, n)
        ref = orc.rotate_axis(ref, d, lo, hi, axis, 0.5 * b * n * dt)
        assert np.abs(mag - ref).max() < 1e-4
    psi = orc.gaussian_3d(d, lo, hi, (2.0, 0, 0), (1.5, 1.5, 1.5), (0, 0.3, 0))
    n0 = orc.norm_sq(psi, 0.75 ** 3)
    out = orc.magnetic_evolve(v, d, lo, hi, 0.01, 0.6, 2, psi, 60)
    assert abs(orc.norm_sq(out, 0.75 ** 3) - n0) < 1e-8


@needs_ref
def test_rotation_and_magnetic_bitwise_equal_reference(orc):
    d, lo, hi = (16, 16, 8), (-8.0, -8.0, -4.0), (8.0, 8.0, 4.0)
    g = orc.gaussian_3d(d, lo, hi, (1.0, 0.5, 0), (1, 1.2, 1), (0.3, 0, 0)).reshape(d[::-1])
    for ax in (0, 1, 2):
        assert np.array_equal(orc.rotate_axis(g, d, lo, hi, ax, 0.3),
                              orc.Ref.rotate_axis(g, d, lo, hi, ax, 0.3))
    v = orc.harmonic_3d(d, lo, hi, 0.5, (0, 0, 0))
    for ax in (2, 0):
        mine = orc.magnetic_evolve(v, d, lo, hi, 0.02, 0.8, ax, g, 7)
        theirs, veff = orc.Ref.magnetic_evolve(v, d, lo, hi, 0.02, 0.8, ax, g, 7)
        assert np.array_equal(mine, theirs)
        assert np.array_equal(orc.magnetic_veff(v, d, lo, hi, 0.8, ax), veff)
