"""The oracle's imaginary-time restatement (oracle/tdse_oracle.c) pinned by the
reference's own tests and, where the reference core is built, bit for bit.

Closed forms restated from tests/imaginary_time_test.cpp (1D harmonic ground
state, soft-Coulomb bound state, normalisation) and tests/deflation_test.cpp
(inner product identities, the first excited harmonic state by deflation),
with the reference's tolerances.
"""
import math
import os

import numpy as np
import pytest

needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                    "libsesref.so")),
    reason="reference core not built here (oracle/Makefile ref needs the reference sources)")

LO, HI, N = -16.0, 16.0, 512  # imaginary_time_test.cpp:21


def sigma_x(psi, lo, hi):
    x = lo + np.arange(psi.size) * (hi - lo) / psi.size
    w = np.abs(psi) ** 2
    m = (w * x).sum() / w.sum()
    return math.sqrt((w * x * x).sum() / w.sum() - m * m)


def test_relaxed_state_is_normalized(orc):
    v = orc.harmonic_1d(LO, HI, N, 1.0, 0.0)
    psi0 = orc.gaussian_1d(LO, HI, N, 1.0, 2.0, 0.0)
    psi, _ = orc.it1d_relax(v, LO, HI, 0.005, psi0, 100)
    assert abs(orc.norm_sq(psi, (HI - LO) / N) - 1.0) < 1e-12


def test_finds_harmonic_ground_state(orc):
    omega = 1.0
    v = orc.harmonic_1d(LO, HI, N, omega, 0.0)
    psi0 = orc.gaussian_1d(LO, HI, N, 1.0, 2.0, 0.0)  # deliberately wrong start
    psi, last = orc.it1d_relax(v, LO, HI, 0.005, psi0, 6000)
    assert abs(orc.mean_energy_1d(psi, LO, HI, v) - 0.5 * omega) < 1e-5
    assert abs(sigma_x(psi, LO, HI) - 1.0 / math.sqrt(2.0 * omega)) < 1e-5
    assert abs(orc.mean_position_1d(psi, LO, HI)) < 1e-8
    # the norm lost by the last (unnormalised) step is exp(-2 E dtau)
    assert abs(-math.log(last) / (2 * 0.005) - 0.5 * omega) < 1e-4


def test_finds_soft_coulomb_bound_state(orc):
    v = orc.soft_coulomb_1d(LO, HI, N, 1.0, 1.0, 0.0)
    psi0 = orc.gaussian_1d(LO, HI, N, 0.5, 1.5, 0.0)
    psi, _ = orc.it1d_relax(v, LO, HI, 0.005, psi0, 10000)
    e0 = orc.mean_energy_1d(psi, LO, HI, v)
    assert e0 < -0.5
    assert abs(e0 - (-0.6698)) < 1e-3  # literature value
    assert abs(orc.mean_position_1d(psi, LO, HI)) < 1e-6


def test_soft_coulomb_formula(orc):
    v = orc.soft_coulomb_1d(-2.0, 2.0, 4, 2.0, 0.5, 0.5)
    x = np.array([-2.0, -1.0, 0.0, 1.0]) - 0.5
    assert np.allclose(v, -2.0 / np.sqrt(x * x + 0.25), rtol=0, atol=1e-15)
    d, lo, hi = (4, 2, 2), (-2.0, -1.0, -1.0), (2.0, 1.0, 1.0)
    v3 = orc.soft_coulomb_3d(d, lo, hi, 1.0, 1.0, (0.0, 0.0, 0.0)).reshape(2, 2, 4)
    assert v3[1, 1, 2] == -1.0  # r = 0
    assert abs(v3[0, 0, 0] - (-1.0 / math.sqrt(4 + 1 + 1 + 1))) < 1e-15


# ---- inner product and deflation (tests/deflation_test.cpp) ------------------

def test_inner_product_exact_single_cell(orc):
    d = (2, 2, 2)  # axis {0, 2, 2}: dV = 1
    a = np.zeros(8, np.complex128)
    b = np.zeros(8, np.complex128)
    f = 1 + 2 * (0 + 2 * 1)  # (i, j, k) = (1, 0, 1)
    a[f] = 1 + 2j
    b[f] = 3 - 1j
    assert orc.inner_product(a, b, 1.0) == complex(1.0, -7.0)
    del d


def test_inner_product_self_and_conjugate_symmetry(orc):
    d, lo, hi = (16, 16, 16), (-8.0,) * 3, (8.0,) * 3
    dv = 1.0
    a = orc.gaussian_3d(d, lo, hi, (1.0, 0.0, 0.0), (1.5, 1.5, 1.5), (0.4, 0.0, 0.0))
    b = orc.gaussian_3d(d, lo, hi, (-0.5, 0.5, 0.0), (1.0, 1.2, 1.4), (0.0, 0.3, 0.0))
    aa = orc.inner_product(a, a, dv)
    assert abs(aa.real - orc.norm_sq(a, dv)) < 1e-12 and abs(aa.imag) < 1e-12
    ab, ba = orc.inner_product(a, b, dv), orc.inner_product(b, a, dv)
    assert abs(ab.real - ba.real) < 1e-12 and abs(ab.imag + ba.imag) < 1e-12


def test_inner_product_even_odd_orthogonal(orc):
    n = 32
    x = -8.0 + np.arange(n) * 16.0 / n
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    env = np.exp(-(xx * xx + y * y + z * z) / 4.0)
    ip = orc.inner_product(env.astype(np.complex128), (xx * env).astype(np.complex128),
                           (16.0 / n) ** 3)
    assert abs(ip.real) < 1e-12 and abs(ip.imag) < 1e-12


def test_deflation_finds_first_excited_harmonic_state(orc):
    omega = 1.0
    d, lo, hi = (32, 32, 32), (-8.0,) * 3, (8.0,) * 3
    dv = 0.5 ** 3
    v = orc.harmonic_3d(d, lo, hi, omega, (0.0, 0.0, 0.0))
    ground0 = orc.gaussian_3d(d, lo, hi, (0, 0, 0), (1.2, 1.2, 1.2), (0, 0, 0))
    ground, _ = orc.it3d_relax(v, d, lo, hi, 0.05, ground0, 600)
    assert abs(orc.mean_energy_3d(ground, d, lo, hi, v) - 1.5 * omega) < 2e-2
    guess = orc.gaussian_3d(d, lo, hi, (1.0, 0.5, 0.0), (1.2, 1.2, 1.2), (0, 0, 0))
    undeflated, _ = orc.it3d_relax(v, d, lo, hi, 0.05, guess, 600)
    assert abs(orc.mean_energy_3d(undeflated, d, lo, hi, v) - 1.5 * omega) < 2e-2
    excited, _ = orc.it3d_relax(v, d, lo, hi, 0.05, guess, 600, lower=[ground])
    assert abs(orc.mean_energy_3d(excited, d, lo, hi, v) - 2.5 * omega) < 2e-2
    ov = orc.inner_product(ground, excited, dv)
    assert abs(ov.real) < 1e-6 and abs(ov.imag) < 1e-6
    assert abs(orc.norm_sq(excited, dv) - 1.0) < 1e-12


def test_relax_any_size_matches_pow2_path(orc):
    """The mixed-radix relax (no reference counterpart) equals the radix-2 one
    to round-off on a power-of-two grid."""
    d, lo, hi = (16, 8, 8), (-6.0, -4.0, -4.0), (6.0, 4.0, 4.0)
    v = orc.harmonic_3d(d, lo, hi, 0.8, (0.3, 0.0, 0.0))
    g = orc.gaussian_3d(d, lo, hi, (0.5, 0.2, 0.0), (1.1, 1.0, 0.9), (0.3, 0.0, 0.0))
    a, na = orc.it3d_relax(v, d, lo, hi, 0.02, g, 30)
    b, nb = orc.it3d_relax(v, d, lo, hi, 0.02, g, 30, any_size=True)
    assert np.abs(a - b).max() < 1e-13 and abs(na - nb) < 1e-13


@needs_ref
def test_relax_restatement_bitwise_equals_reference(orc):
    R = orc.Ref
    v = orc.soft_coulomb_1d(LO, HI, N, 1.0, 1.0, 0.0)
    assert np.array_equal(v, R.soft_coulomb_1d(LO, HI, N, 1.0, 1.0, 0.0))
    psi0 = orc.gaussian_1d(LO, HI, N, 0.5, 1.5, 0.0)
    a, _ = orc.it1d_relax(v, LO, HI, 0.005, psi0, 300)
    assert np.array_equal(a, R.it1d_relax(v, LO, HI, 0.005, psi0, 300))
    d, lo, hi = (16, 8, 8), (-8.0, -4.0, -4.0), (8.0, 4.0, 4.0)
    v3 = orc.harmonic_3d(d, lo, hi, 1.0, (0.0, 0.0, 0.0))
    assert np.array_equal(orc.soft_coulomb_3d(d, lo, hi, 1.2, 0.7, (0.1, 0, 0)),
                          R.soft_coulomb_3d(d, lo, hi, 1.2, 0.7, (0.1, 0, 0)))
    for mine, theirs in zip(orc.it3d_tables(v3, d, lo, hi, 0.05, 0.7),
                            R.it3d_tables(v3, d, lo, hi, 0.05, 0.7)):
        assert np.array_equal(mine, theirs)
    g = orc.gaussian_3d(d, lo, hi, (0.5, 0.2, 0), (1.2, 1.2, 1.2), (0, 0, 0)).reshape(d[::-1])
    x, _ = orc.it3d_relax(v3, d, lo, hi, 0.05, g, 50)
    y = R.it3d_relax(v3, d, lo, hi, 0.05, g, 50)
    assert np.array_equal(x, y)
    g2 = orc.gaussian_3d(d, lo, hi, (1.0, 0.5, 0), (1.2, 1.2, 1.2), (0, 0, 0)).reshape(d[::-1])
    x2, _ = orc.it3d_relax(v3, d, lo, hi, 0.05, g2, 40, lower=[x])
    assert np.array_equal(x2, R.it3d_relax(v3, d, lo, hi, 0.05, g2, 40, lower=[y]))
    assert orc.inner_product(x, g2, 1.0) == R.inner_product(x, g2, d, lo, hi)


def test_width_and_momentum_closed_forms(orc):
    """A Gaussian packet of density width s and wavenumber k0 has sigma = s
    and <k> = k0 (resolved grid): pins orc_sigma_* and orc_mean_momentum_*."""
    p = orc.gaussian_1d(-16.0, 16.0, 512, 1.3, 1.7, 0.6)
    assert abs(orc.sigma_x_1d(p, -16.0, 16.0) - 1.7) < 1e-9
    assert abs(orc.mean_momentum_1d(p, -16.0, 16.0) - 0.6) < 1e-9
    d, lo, hi = (32, 32, 32), (-10.0,) * 3, (10.0,) * 3
    g = orc.gaussian_3d(d, lo, hi, (0.5, 0.2, 0.0), (1.6, 1.4, 1.2), (0.4, -0.3, 0.2))
    np.testing.assert_allclose(orc.sigma_position_3d(g, d, lo, hi), (1.6, 1.4, 1.2), atol=1e-6)
    np.testing.assert_allclose(orc.mean_momentum_3d(g, d, lo, hi), (0.4, -0.3, 0.2), atol=1e-6)


@needs_ref
def test_width_and_momentum_bitwise_equal_reference(orc):
    R = orc.Ref
    p = orc.gaussian_1d(-16.0, 16.0, 512, 1.3, 1.7, 0.6)
    assert orc.sigma_x_1d(p, -16.0, 16.0) == R.sigma_x_1d(p, -16.0, 16.0)
    assert orc.mean_momentum_1d(p, -16.0, 16.0) == R.mean_momentum_1d(p, -16.0, 16.0)
    d, lo, hi = (16, 8, 8), (-8.0, -4.0, -4.0), (8.0, 4.0, 4.0)
    g = orc.gaussian_3d(d, lo, hi, (0.5, 0.2, 0.0), (1.2, 1.1, 1.0), (0.4, -0.3, 0.2))
    assert orc.sigma_position_3d(g, d, lo, hi) == R.sigma_position_3d(g, d, lo, hi)
    assert orc.mean_momentum_3d(g, d, lo, hi) == R.mean_momentum_3d(g, d, lo, hi)
