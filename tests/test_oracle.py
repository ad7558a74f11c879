"""Pins the oracle (oracle/tdse_oracle.c) before it is trusted as the checker.

Three kinds of pins, chosen so that a plausible slip (wrong kernel sign, a
transposed axis, a dropped 1/N, a wrong table formula, a swapped mask order)
fails at least one:

1. The closed forms of the reference's own tests (tests/fft_test.cpp,
   tests/fft3_test.cpp, tests/propagator_test.cpp,
   tests/propagator_tables_test.cpp, tests/potential_test.cpp,
   tests/harmonic_dynamics_test.cpp), restated here with the same tolerances.
2. Bit-for-bit equality with the reference core compiled from its own sources
   (oracle/_ref, built by oracle/Makefile) where that library is present.
3. Golden vectors written by tests/golden/make_golden.py from that reference
   build (tests/golden/*.npz), which travel with the repository.
Mixed-radix transforms (no reference counterpart) are pinned by a brute-force
DFT, numpy.fft and Parseval/inverse/linearity.
"""
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def det_signal(n):
    """tests/fft_test.cpp deterministic_signal (no RNG)."""
    x = np.arange(n, dtype=np.float64)
    return np.sin(1.3 * x) + 0.1 * x + 1j * (np.cos(2.7 * x) - 0.05 * x)


# ---- FFT closed forms (tests/fft_test.cpp, tests/fft3_test.cpp) -------------

def test_fft_size_one_is_identity(orc):
    assert orc.fft1d(np.array([3 - 4j]))[0] == 3 - 4j


def test_fft_dc_and_delta(orc):
    a = orc.fft1d(np.ones(8, complex))
    assert abs(a[0] - 8) < 1e-12 and np.abs(a[1:]).max() < 1e-12
    d = np.zeros(8, complex)
    d[0] = 1
    assert np.abs(orc.fft1d(d) - 1).max() < 1e-12


def test_fft_kernel_sign(orc):
    n, k0 = 16, 3
    i = np.arange(n)
    a = orc.fft1d(np.exp(2j * np.pi * k0 * i / n))
    want = np.zeros(n, complex)
    want[k0] = 16
    assert np.abs(a - want).max() < 1e-12  # e^{+} kernel would spike at 13


def test_fft_cosine_pair(orc):
    n, k0 = 16, 3
    a = orc.fft1d(np.cos(2 * np.pi * k0 * np.arange(n) / n) + 0j)
    want = np.zeros(n, complex)
    want[k0] = want[n - k0] = 8
    assert np.abs(a - want).max() < 1e-12


def test_fft_inverse_linear_parseval(orc):
    x = det_signal(16)
    assert np.abs(orc.ifft1d(orc.fft1d(x)) - x).max() < 1e-12
    x8 = det_signal(8)
    y8 = 0.5 * np.arange(8) - 1j
    assert np.abs(orc.fft1d(2 * x8 + y8) - (2 * orc.fft1d(x8) + orc.fft1d(y8))).max() < 1e-12
    assert abs(np.sum(abs(x) ** 2) - np.sum(abs(orc.fft1d(x)) ** 2) / 16) < 1e-10


def test_fft_rejects_non_pow2(orc):
    with pytest.raises(ValueError):
        orc.fft1d(np.ones(12, complex))


def test_fft_matches_numpy(orc):
    rng = np.random.default_rng(5)
    for n in (2, 4, 32, 1024, 65536):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert np.abs(orc.fft1d(x) - np.fft.fft(x)).max() < 1e-11 * max(1, math.log2(n))
        assert np.abs(orc.ifft1d(x) - np.fft.ifft(x)).max() < 1e-13 * max(1, math.log2(n))


def test_fft3_dc_plane_wave_inverse_parseval(orc):
    f = orc.fft3d(np.ones((2, 4, 8), complex))
    assert abs(f[0, 0, 0] - 64) < 1e-12 and abs(f[0, 0, 1]) < 1e-12 and abs(f[1, 0, 0]) < 1e-12
    nx, ny, nz = 8, 16, 4
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    th = 2 * np.pi * (3.0 * i / nx + 5.0 * j / ny + 1.0 * k / nz)
    f = orc.fft3d(np.exp(1j * th))
    want = np.zeros((nz, ny, nx), complex)
    want[1, 5, 3] = 512  # (i, j, k) = (3, 5, 1): a transposed axis would miss
    assert np.abs(f - want).max() < 1e-9
    t = np.arange(64, dtype=float)
    x = (np.sin(1.3 * t) + 0.1 * t + 1j * np.cos(2.7 * t)).reshape(2, 4, 8)
    assert np.abs(orc.ifft3d(orc.fft3d(x)) - x).max() < 1e-12
    y = (np.cos(0.9 * t) + 0.2j * np.sin(1.7 * t)).reshape(2, 8, 4)
    assert abs(np.sum(abs(y) ** 2) - np.sum(abs(orc.fft3d(y)) ** 2) / 64) < 1e-10
    assert np.abs(orc.fft3d(y) - np.fft.fftn(y)).max() < 1e-12


# ---- mixed radix (no reference counterpart) ----------------------------------

@pytest.mark.parametrize("n", [3, 5, 6, 10, 12, 48, 96, 160, 192, 7, 14])
def test_mixed_radix_matches_brute_and_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    brute = orc.dft_brute(x)
    assert np.abs(brute - np.fft.fft(x)).max() < 1e-12 * n
    assert np.abs(orc.fft_any(x) - brute).max() < 1e-12 * n
    assert np.abs(orc.fft_any(orc.fft_any(x), +1) / n - x).max() < 1e-13 * n


def test_mixed_radix_3d_matches_numpy(orc):
    rng = np.random.default_rng(9)
    x = rng.standard_normal((4, 160, 96)) + 1j * rng.standard_normal((4, 160, 96))
    assert np.abs(orc.fft3d_any(x) - np.fft.fftn(x)).max() < 1e-10
    assert np.abs(orc.fft3d_any(x, +1) - np.fft.ifftn(x) * x.size).max() < 1e-10


# ---- tables (tests/propagator_tables_test.cpp) --------------------------------

def test_tables_equal_defining_formula(orc):
    dims, lo, hi = (8, 4, 2), (-2.0, -1.0, -0.5), (2.0, 1.0, 0.5)
    v = 0.01 * np.arange(64, dtype=float)
    hv, kin = orc.so3d_tables(v, dims, lo, hi, 0.05, 1.0)
    assert np.array_equal(hv, np.cos(-0.5 * v * 0.05) + 1j * np.sin(-0.5 * v * 0.05))
    kx = orc.wavenumbers(8, -2.0, 2.0)
    ky = orc.wavenumbers(4, -1.0, 1.0)
    kz = orc.wavenumbers(2, -0.5, 0.5)
    k, j, i = np.meshgrid(np.arange(2), np.arange(4), np.arange(8), indexing="ij")
    th = -0.5 * (kx[i] ** 2 + ky[j] ** 2 + kz[k] ** 2) / 1.0 * 0.05
    assert np.array_equal(kin.reshape(2, 4, 8), np.cos(th) + 1j * np.sin(th))


def test_wavenumbers_fftfreq_layout(orc):
    """tests/spectral_test.cpp: fftfreq order, DC at n = 1."""
    k = orc.wavenumbers(8, 0.0, 2 * np.pi)
    assert np.array_equal(k, np.array([0, 1, 2, 3, -4, -3, -2, -1], float))
    assert orc.wavenumbers(1, 0.0, 1.0)[0] == 0.0
    assert np.allclose(orc.wavenumbers(16, -3, 5), 2 * np.pi * np.fft.fftfreq(16, 8 / 16))


# ---- propagation (tests/propagator_test.cpp, harmonic_dynamics_test.cpp) -----

GRID = (-40.0, 40.0, 1024)


def free_run(orc, steps=200, dt=0.05, x0=-15.0, s0=1.0, k0=2.0):
    lo, hi, n = GRID
    psi = orc.gaussian_1d(lo, hi, n, x0, s0, k0)
    return orc.so1d_evolve(np.zeros(n), lo, hi, dt, psi, steps)


def test_free_packet_norm_velocity_dispersion_momentum(orc):
    lo, hi, n = GRID
    psi = free_run(orc)
    h = (hi - lo) / n
    x = lo + np.arange(n) * h
    w = abs(psi) ** 2
    assert abs(orc.norm_sq(psi, h) - 1.0) < 1e-12
    t = 0.05 * 200
    mean = orc.mean_position_1d(psi, lo, hi)
    assert abs(mean - (-15.0 + 2.0 * t)) < 1e-6
    sigma = math.sqrt(max(0.0, np.sum(x * x * w) / np.sum(w) - mean ** 2))
    assert abs(sigma - math.sqrt(1 + (t / 2) ** 2)) < 1e-6
    phi = orc.fft1d(psi)
    k = orc.wavenumbers(n, lo, hi)
    assert abs(np.sum(k * abs(phi) ** 2) / np.sum(abs(phi) ** 2) - 2.0) < 1e-8


def test_zero_steps_identity(orc):
    lo, hi, n = GRID
    psi = orc.gaussian_1d(lo, hi, n, -15.0, 1.0, 2.0)
    assert np.array_equal(orc.so1d_evolve(np.zeros(n), lo, hi, 0.05, psi, 0), psi)


def test_harmonic_ehrenfest(orc):
    """Quadratic V: <x> follows the exact Stormer-Verlet map of the split step."""
    lo, hi, n = GRID
    w = 0.5
    v = orc.harmonic_1d(lo, hi, n, w)
    psi0 = orc.gaussian_1d(lo, hi, n, 3.0, 1 / math.sqrt(2 * w), 0.0)
    dt, steps = 0.05, 100
    psi = orc.so1d_evolve(v, lo, hi, dt, psi0, steps)
    x, p = 3.0, 0.0  # half kick, drift, half kick on (x, p) is linear
    for _ in range(steps):
        p -= 0.5 * dt * w * w * x
        x += dt * p
        p -= 0.5 * dt * w * w * x
    assert abs(orc.mean_position_1d(psi, lo, hi) - x) < 1e-8


HGRID = (-16.0, 16.0, 512)  # tests/harmonic_dynamics_test.cpp:25-31


def coherent(orc):
    lo, hi, n = HGRID
    return orc.gaussian_1d(lo, hi, n, 2.0, 1.0 / math.sqrt(2.0), 0.0)


def test_harmonic_dynamics_reference_cases(orc):
    """tests/harmonic_dynamics_test.cpp: energy 2.5, mirror point at T/2,
    no spreading, momentum at T/4, energy over T, Strang O(dt^2)."""
    lo, hi, n = HGRID
    h = (hi - lo) / n
    x = lo + np.arange(n) * h
    v = orc.harmonic_1d(lo, hi, n, 1.0)
    dt = 2.0 * math.pi / 1000.0
    psi0 = coherent(orc)
    assert abs(orc.mean_energy_1d(psi0, lo, hi, v) - 2.5) < 1e-8
    half = orc.so1d_evolve(v, lo, hi, dt, psi0, 500)
    assert abs(orc.mean_position_1d(half, lo, hi) + 2.0) < 1e-5
    assert abs(orc.norm_sq(half, h) - 1.0) < 1e-12
    w = abs(half) ** 2
    m = np.sum(x * w) / np.sum(w)
    assert abs(math.sqrt(np.sum(x * x * w) / np.sum(w) - m * m) - 1 / math.sqrt(2)) < 1e-5
    quarter = orc.so1d_evolve(v, lo, hi, dt, psi0, 250)
    phi = orc.fft1d(quarter)
    k = orc.wavenumbers(n, lo, hi)
    assert abs(np.sum(k * abs(phi) ** 2) / np.sum(abs(phi) ** 2) + 2.0) < 1e-4
    assert abs(orc.mean_position_1d(quarter, lo, hi)) < 1e-4
    full = orc.so1d_evolve(v, lo, hi, dt, psi0, 1000)
    assert abs(orc.mean_energy_1d(full, lo, hi, v) - 2.5) < 1e-4
    exact = 2.0 * math.cos(1.0)
    e_coarse = abs(orc.mean_position_1d(orc.so1d_evolve(v, lo, hi, 0.1, psi0, 10), lo, hi) - exact)
    e_fine = abs(orc.mean_position_1d(orc.so1d_evolve(v, lo, hi, 0.05, psi0, 20), lo, hi) - exact)
    assert e_coarse > 1e-9
    assert 3.4 < e_coarse / e_fine < 4.6


def test_3d_free_dispersion_exact_at_dt1(orc):
    """dynamics3_test.cpp: V = 0 makes the split step exact at any dt."""
    dims, lo, hi = (32, 32, 16), (-16, -16, -8), (16, 16, 8)
    psi0 = orc.gaussian_3d(dims, lo, hi, (0, 0, 0), (1.0, 1.5, 0.8), (0.5, 0, 0)).reshape(16, 32, 32)
    one = orc.so3d_evolve(np.zeros(32 * 32 * 16), dims, lo, hi, 1.0, psi0, 1)
    many = orc.so3d_evolve(np.zeros(32 * 32 * 16), dims, lo, hi, 0.1, psi0, 10)
    assert np.abs(one - many).max() < 1e-12


def test_mass_drift(orc):
    """MassParameter.FreePacketDriftsAtK0OverM (propagator_test.cpp)."""
    dims, lo, hi = (128, 1, 1), (-32.0, -1.0, -1.0), (32.0, 1.0, 1.0)
    psi0 = orc.gaussian_3d(dims, lo, hi, (-8, 0, 0), (2.0, 0.5, 0.5), (1.0, 0, 0)).reshape(1, 1, 128)
    psi = orc.so3d_evolve(np.zeros(128), dims, lo, hi, 4.0 / 200, psi0, 200, mass=2.0)
    assert abs(orc.mean_position_3d(psi, dims, lo, hi)[0] - (-8 + 0.5 * 4.0)) < 0.05


def test_mask_formula_and_order(orc):
    """potential_test.cpp mask formula; the mask multiplies after each step."""
    lo, hi, n = -10.0, 10.0, 64
    m = orc.absorbing_mask_1d(lo, hi, n, 4.0)
    x = lo + np.arange(n) * (hi - lo) / n
    d = np.minimum(x - lo, hi - x)
    want = np.where(d >= 4.0, 1.0, np.sin(0.5 * np.pi * d / 4.0) ** 2)
    assert np.abs(m - want).max() < 1e-15 and m[0] == 0.0
    assert orc.absorbing_mask_1d(lo, hi, 1, 4.0)[0] == 1.0
    v = orc.harmonic_1d(lo, hi, n, 0.3)
    psi0 = orc.gaussian_1d(lo, hi, n, 2.0, 1.0, 1.0)
    one = orc.so1d_evolve(v, lo, hi, 0.1, psi0, 1) * m
    assert np.array_equal(orc.so1d_evolve(v, lo, hi, 0.1, psi0, 1, mask=m), one)


def test_barrier_half_open_and_coulomb(orc):
    lo, hi, n = -2.0, 2.0, 8  # coords -2, -1.5, ..., 1.5
    v = orc.barrier_1d(lo, hi, n, 1.5, -1.0, 0.5)
    assert list(v) == [0, 0, 1.5, 1.5, 1.5, 0, 0, 0]
    dims, l3, h3 = (8, 8, 8), (-2.0,) * 3, (2.0,) * 3
    c = orc.regularized_coulomb_3d(dims, l3, h3, 1.0, (0, 0, 0))
    f = orc.coulomb_floor_3d(dims, l3, h3, 1.0, (0, 0, 0))
    center = 4 + 8 * (4 + 8 * 4)
    assert c[center] == -1.0 * 2.3800774 / 0.5 and f[center] == -1.0 / 0.25
    assert c[center + 1] == f[center + 1] == -1.0 / 0.5


def test_slab_probability_gaussian_tail(orc):
    """P(x >= a) of a normalised Gaussian: the half-open sum over grid points is
    the midpoint rule from the first included point minus h/2, i.e.
    erfc((x_first - h/2 - x0) / (sqrt(2) sigma)) / 2."""
    dims, lo, hi = (256, 8, 8), (-20.0, -8.0, -8.0), (20.0, 8.0, 8.0)
    psi = orc.gaussian_3d(dims, lo, hi, (0.3, 0, 0), (1.2, 1.0, 1.0), (0, 0, 0))
    h = 40.0 / 256
    for a in (1.0, 0.3, -2.0):
        first = lo[0] + math.ceil((a - lo[0]) / h - 1e-12) * h
        want = 0.5 * math.erfc((first - 0.5 * h - 0.3) / (math.sqrt(2) * 1.2))
        got = orc.probability_in_slab_3d(psi, dims, lo, hi, 0, a, 1e300)
        assert abs(got - want) < 2e-4
    assert abs(orc.probability_in_slab_3d(psi, dims, lo, hi, 0, -1e300, 1e300) - 1) < 1e-12


def test_box_probability_pins(orc):
    """The box: one restricted axis is the slab bit for bit; for a separable
    (product) state the sum over a box factorises, so box * N^2 equals the
    product of the three slabs, N the whole grid's sum (the norm), up to the
    rounding of sequential sums over millions of cells; the Gaussian closed
    form per axis as in the slab test; an inverted box is empty; axes of
    extent 1 are not restricted."""
    dims, lo, hi = (256, 192, 128), (-12.0, -12.0, -8.0), (12.0, 12.0, 8.0)
    x0, sig = (0.3, -0.7, 0.2), (1.4, 1.6, 1.5)
    psi = orc.gaussian_3d(dims, lo, hi, x0, sig, (0.4, 0.0, -0.3))
    inf = 1e300
    a, b = (-1.0, -2.5, -0.5), (2.0, 0.5, 1.5)
    for ax in range(3):
        lo_b, hi_b = [-inf] * 3, [inf] * 3
        lo_b[ax], hi_b[ax] = a[ax], b[ax]
        assert (orc.probability_in_box_3d(psi, dims, lo, hi, lo_b, hi_b)
                == orc.probability_in_slab_3d(psi, dims, lo, hi, ax, a[ax], b[ax]))
    box = orc.probability_in_box_3d(psi, dims, lo, hi, a, b)
    norm = orc.probability_in_box_3d(psi, dims, lo, hi, [-inf] * 3, [inf] * 3)
    assert norm == orc.probability_in_slab_3d(psi, dims, lo, hi, 0, -inf, inf)
    assert abs(norm - 1) < 1e-10
    slabs = [orc.probability_in_slab_3d(psi, dims, lo, hi, ax, a[ax], b[ax]) for ax in range(3)]
    assert abs(box * norm ** 2 - slabs[0] * slabs[1] * slabs[2]) < 1e-10
    want = 1.0
    for ax in range(3):
        h = (hi[ax] - lo[ax]) / dims[ax]
        first = lo[ax] + math.ceil((a[ax] - lo[ax]) / h - 1e-12) * h
        past = lo[ax] + math.ceil((b[ax] - lo[ax]) / h - 1e-12) * h
        s2 = math.sqrt(2) * sig[ax]
        want *= 0.5 * (math.erfc((first - 0.5 * h - x0[ax]) / s2)
                       - math.erfc((past - 0.5 * h - x0[ax]) / s2))
    assert abs(box - want) < 5e-4
    assert orc.probability_in_box_3d(psi, dims, lo, hi, b, a) == 0.0
    d2, lo2, hi2 = (64, 48, 1), (-16.0, -12.0, 0.0), (16.0, 12.0, 1.0)
    psi2 = orc.gaussian_3d(d2, lo2, hi2, (0.5, -0.5, 0.0), (1.5, 1.2, 1.0), (0.0, 0.0, 0.0))
    z_excluded = orc.probability_in_box_3d(psi2, d2, lo2, hi2, (-1, -1, 5.0), (2, 1, 6.0))
    z_open = orc.probability_in_box_3d(psi2, d2, lo2, hi2, (-1, -1, -inf), (2, 1, inf))
    assert z_excluded == z_open > 0


# ---- bitwise against the reference build -------------------------------------

needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                    "libsesref.so")),
    reason="reference core not built here (oracle/Makefile ref needs the reference sources)")


@needs_ref
def test_restatement_bitwise_equals_reference(orc):
    R = orc.Ref
    rng = np.random.default_rng(1)
    for n in (1, 2, 8, 1024, 65536):
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert np.array_equal(orc.fft1d(a), R.fft1d(a))
        assert np.array_equal(orc.ifft1d(a), R.ifft1d(a))
    a = rng.standard_normal((4, 16, 8)) + 1j * rng.standard_normal((4, 16, 8))
    assert np.array_equal(orc.fft3d(a), R.fft3d(a))
    assert np.array_equal(orc.ifft3d(a), R.ifft3d(a))
    dims, lo, hi = (16, 8, 32), (-5.0, -4.0, -8.0), (5.0, 4.0, 8.0)
    v = orc.regularized_coulomb_3d(dims, lo, hi, 1.0, (0, 0, 0))
    assert np.array_equal(v, R.regularized_coulomb_3d(dims, lo, hi, 1.0, (0, 0, 0)))
    m = orc.absorbing_mask_3d(dims, lo, hi, 2.0)
    assert np.array_equal(m, R.absorbing_mask_3d(dims, lo, hi, 2.0))
    psi = orc.gaussian_3d(dims, lo, hi, (0.5, 0, 1), (1.0, 1.2, 2), (0.7, 0, -0.3))
    assert np.array_equal(psi, R.gaussian_3d(dims, lo, hi, (0.5, 0, 1), (1.0, 1.2, 2), (0.7, 0, -0.3)))
    h1, k1 = orc.so3d_tables(v, dims, lo, hi, 0.01, 0.38)
    h2, k2 = R.so3d_tables(v, dims, lo, hi, 0.01, 0.38)
    assert np.array_equal(h1, h2) and np.array_equal(k1, k2)
    psi = psi.reshape(dims[::-1])
    assert np.array_equal(orc.so3d_evolve(v, dims, lo, hi, 0.01, psi, 7, mask=m),
                          R.so3d_evolve(v, dims, lo, hi, 0.01, psi, 7, mask=m))
    lo1, hi1, n1 = GRID
    v1 = orc.harmonic_1d(lo1, hi1, n1, 0.1)
    p1 = orc.gaussian_1d(lo1, hi1, n1, -15.0, 1.0, 2.0)
    assert np.array_equal(orc.so1d_evolve(v1, lo1, hi1, 0.05, p1, 20),
                          R.so1d_evolve(v1, lo1, hi1, 0.05, p1, 20))
    assert orc.mean_energy_1d(p1, lo1, hi1, v1) == R.mean_energy_1d(p1, lo1, hi1, v1)
    assert orc.mean_energy_3d(psi, dims, lo, hi, v, 0.38) == R.mean_energy_3d(psi, dims, lo, hi, v, 0.38)


# ---- golden vectors from the reference build ---------------------------------

def golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    return np.load(path, allow_pickle=False)


def test_golden_fft(orc):
    g = golden("fft.npz")
    assert np.array_equal(orc.fft1d(g["x1024"]), g["fft1024"])
    assert np.array_equal(orc.ifft1d(g["x1024"]), g["ifft1024"])
    assert np.array_equal(orc.fft3d(g["x3"]), g["fft3"])
    assert np.array_equal(orc.ifft3d(g["x3"]), g["ifft3"])


def test_golden_step_small(orc):
    g = golden("step_small.npz")
    dims, lo, hi = tuple(g["dims"]), tuple(g["lo"]), tuple(g["hi"])
    out = orc.so3d_evolve(g["v"], dims, lo, hi, float(g["dt"]), g["psi0"], int(g["steps"]),
                          mass=float(g["mass"]), mask=g["mask"])
    assert np.array_equal(out, g["psi"])
    o1 = orc.so1d_evolve(g["v1"], float(g["lo1"]), float(g["hi1"]), float(g["dt1"]), g["psi1_0"],
                         int(g["steps1"]), mask=g["mask1"])
    assert np.array_equal(o1, g["psi1"])


def test_energy_variance_restatement(orc):
    """energy_variance (core/src/observables.ixx:92-119): the C restatement
    against the reference core where it is built here (bit for bit), and the
    closed forms of tests/observables_test.cpp:55-93 and 113-121: zero on the
    oscillator's eigenstates, (E1 - E0)^2 / 4 on their equal mix, scale
    invariance, positive on a displaced packet, and round-off level on an
    eigenstate of energy 10002."""
    from ho_states import ho_eigenstate
    omega, xmin, xmax, n = 0.25, -20.0, 20.0, 256
    v = orc.harmonic_1d(xmin, xmax, n, omega)
    p0 = ho_eigenstate(xmin, xmax, n, omega, 0)
    p1 = ho_eigenstate(xmin, xmax, n, omega, 1)
    mix = p0 + p1
    assert abs(orc.energy_variance_1d(p0, xmin, xmax, v)) < 1e-10
    assert abs(orc.energy_variance_1d(p1, xmin, xmax, v)) < 1e-10
    assert abs(orc.energy_variance_1d(mix, xmin, xmax, v) - omega * omega / 4.0) < 1e-10
    g = orc.gaussian_1d(-16.0, 16.0, 128, 1.0, 1.3, 0.4)
    vg = orc.harmonic_1d(-16.0, 16.0, 128, 0.5)
    var1 = orc.energy_variance_1d(g, -16.0, 16.0, vg)
    assert var1 > 0.0
    assert abs(orc.energy_variance_1d(3.0 * g, -16.0, 16.0, vg) - var1) < 1e-9 * (1.0 + var1)
    if orc.ref_available():
        for psi, a, b, vv in ((p0, xmin, xmax, v), (mix, xmin, xmax, v), (g, -16.0, 16.0, vg)):
            assert orc.energy_variance_1d(psi, a, b, vv) == orc.Ref.energy_variance_1d(psi, a, b, vv)
    # the reference's high-rung lock on the 65536 scene grid
    vh = orc.harmonic_1d(-100.0, 100.0, 65536, 4.0)
    ph = ho_eigenstate(-100.0, 100.0, 65536, 4.0, 2500)
    assert orc.energy_variance_1d(ph, -100.0, 100.0, vh) < 1e-8
    assert abs(orc.mean_energy_1d(ph, -100.0, 100.0, vh) - 2500.5 * 4.0) < 1e-6 * 10002.0
