"""The oracle's dipole-drive restatement (orc_driven_step, core/src/drive.ixx)
pinned by the reference's tests/drive_test.cpp closed forms and, where the
reference core is built, bit for bit."""
import math
import os

import numpy as np
import pytest

needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                    "libsesref.so")),
    reason="reference core not built here (oracle/Makefile ref needs the reference sources)")


def cube(lo, hi, n):
    return (n, n, n), (lo,) * 3, (hi,) * 3


def test_zero_amplitude_matches_static_bitwise(orc):
    d, lo, hi = cube(-8.0, 8.0, 16)
    v = orc.harmonic_3d(d, lo, hi, 1.0, (0, 0, 0))
    psi0 = orc.gaussian_3d(d, lo, hi, (1.0, 0, 0), (1, 1, 1), (0, 0, 0))
    a = orc.driven_evolve(v, d, lo, hi, 0.02, psi0, 10, (0, 0, 1), 0.0, 0.5)
    assert np.array_equal(a, orc.so3d_evolve(v, d, lo, hi, 0.02, psi0, 10))


def test_constant_field_obeys_ehrenfest(orc):
    d, lo, hi = cube(-16.0, 16.0, 32)
    dt, e0, steps = 0.04, 0.05, 50
    v = np.zeros(32 ** 3)
    psi0 = orc.gaussian_3d(d, lo, hi, (0, 0, 0), (1.5, 1.5, 1.5), (0, 0, 0))
    psi = orc.driven_evolve(v, d, lo, hi, dt, psi0, steps, (0, 0, 1), e0, 0.0)
    t = steps * dt
    assert abs(orc.mean_momentum_3d(psi, d, lo, hi)[2] - (-e0 * t)) < 1e-8
    assert abs(orc.mean_position_3d(psi, d, lo, hi)[2] - (-0.5 * e0 * t * t)) < 1e-6
    assert abs(orc.norm_sq(psi, 1.0) - 1.0) < 1e-12


def test_resonant_coherent_ladder_and_selection_rule(orc):
    w0 = 1.0
    d, lo, hi = cube(-8.0, 8.0, 32)
    v = orc.harmonic_3d(d, lo, hi, w0, (0, 0, 0))
    dt = 0.05
    x = -8.0 + np.arange(32) * 0.5
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    env = np.exp(-0.5 * w0 * (x3 ** 2 + y3 ** 2 + z3 ** 2))
    dv = 0.5 ** 3

    def normed(a):
        a = a.astype(np.complex128)
        return a / math.sqrt(orc.norm_sq(a, dv))

    ground, ez, ey = normed(env), normed(z3 * env), normed(y3 * env)
    e0, steps = 0.2, 283
    t = steps * dt
    psi = orc.driven_evolve(v, d, lo, hi, dt, ground, steps, (0, 0, 1), e0, w0)
    expected_z = -(e0 / (2.0 * w0)) * t * math.sin(w0 * t)
    assert abs(orc.mean_position_3d(psi, d, lo, hi)[2] - expected_z) < 0.02
    alpha_sq = (e0 * t / (2.0 * math.sqrt(2.0 * w0))) ** 2
    pop = lambda s: abs(orc.inner_product(s, psi, dv)) ** 2
    assert abs(pop(ground) - math.exp(-alpha_sq)) < 0.06
    assert abs(pop(ez) - alpha_sq * math.exp(-alpha_sq)) < 0.06
    assert pop(ey) < 1e-4
    assert abs(orc.norm_sq(psi, dv) - 1.0) < 1e-10


@needs_ref
def test_driven_step_bitwise_equals_reference(orc):
    d, lo, hi = (16, 8, 8), (-8.0, -4.0, -4.0), (8.0, 4.0, 4.0)
    v = orc.harmonic_3d(d, lo, hi, 1.0, (0, 0, 0))
    g = orc.gaussian_3d(d, lo, hi, (1.0, 0, 0), (1, 1, 1), (0.2, 0, 0)).reshape(d[::-1])
    for axis, amp, om, t0 in [((0.3, 0.2, 1.0), 0.5, 0.7, 0.1), ((1, 0, 0), 0.05, 0.0, 0.0)]:
        a = orc.driven_evolve(v, d, lo, hi, 0.02, g, 10, axis, amp, om, t0=t0, mass=0.8)
        b = orc.Ref.driven_evolve(v, d, lo, hi, 0.02, g, 10, axis, amp, om, t0=t0, mass=0.8)
        assert np.array_equal(a, b)
