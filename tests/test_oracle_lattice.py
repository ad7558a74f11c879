"""The oracle's restatement of PeierlsLattice2D (core/src/lattice2d.ixx),
pinned by the reference's tests/lattice2d_test.cpp closed forms and, where
the reference core is built, bit for bit."""
import cmath
import math
import os

import numpy as np
import pytest

needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                    "libsesref.so")),
    reason="reference core not built here (oracle/Makefile ref needs the reference sources)")


def plane(lx, nx, ly, ny):
    return (nx, ny, 1), (-lx, -ly, -1.0), (lx, ly, 1.0)


def packet(d, lo, hi, x0, y0, sigma, k0):
    x = lo[0] + np.arange(d[0]) * (hi[0] - lo[0]) / d[0]
    y = lo[1] + np.arange(d[1]) * (hi[1] - lo[1]) / d[1]
    yy, xx = np.meshgrid(y, x, indexing="ij")
    env = np.exp(-((xx - x0) ** 2 + (yy - y0) ** 2) / (4 * sigma * sigma))
    psi = (env * np.exp(1j * k0 * xx)).astype(np.complex128).reshape(1, d[1], d[0])
    dv = (hi[0] - lo[0]) / d[0] * (hi[1] - lo[1]) / d[1] * (hi[2] - lo[2])
    return psi / math.sqrt((np.abs(psi) ** 2).sum() * dv), dv


def mean_x(psi, d, lo, hi):
    x = lo[0] + np.arange(d[0]) * (hi[0] - lo[0]) / d[0]
    w = (np.abs(psi.reshape(d[1], d[0])) ** 2)
    return float((w * x[None, :]).sum() / w.sum())


def test_unitary_to_round_off(orc):
    d, lo, hi = plane(10.0, 64, 10.0, 64)
    psi, dv = packet(d, lo, hi, -3.0, 1.0, 2.0, 1.5)
    out, _ = orc.lat2d_evolve(np.zeros(64 * 64), d, lo, hi, 0.02, psi, 500)
    assert abs((np.abs(out) ** 2).sum() * dv - 1.0) < 1e-12


def test_free_packet_moves_at_the_lattice_group_velocity(orc):
    d, lo, hi = plane(30.0, 256, 10.0, 32)
    dt, k0, steps = 0.01, 1.2, 1500
    psi, _ = packet(d, lo, hi, -15.0, 0.0, 3.0, k0)
    x0 = mean_x(psi, d, lo, hi)
    out, _ = orc.lat2d_evolve(np.zeros(256 * 32), d, lo, hi, dt, psi, steps)
    h = 60.0 / 256
    vg = math.sin(k0 * h) / h
    assert abs(mean_x(out, d, lo, hi) - x0 - vg * dt * steps) < 0.01 * vg * dt * steps


def _loop(lx, d, i, j):
    nx = d[0]
    ux = lambda a, b: lx[b * nx + a]
    return ux(i, j) * 1.0 * np.conj(ux(i, j + 1)) * 1.0


def test_solenoid_links_carry_pure_topology(orc):
    d, lo, hi = plane(8.0, 32, 8.0, 32)
    h = 0.5
    phi = 1.9
    xs, ys = lo[0] + 15 * h + 0.5 * h, lo[1] + 12 * h + 0.5 * h
    lx = orc.lat2d_links(d, lo, hi, 2, phi, xs, ys, True)
    hot = 0
    for j in range(31):
        for i in range(31):
            flux = cmath.phase(_loop(lx, d, i, j))
            if i == 15 and j == 12:
                assert abs(math.remainder(flux - phi, 2 * math.pi)) < 1e-12
                hot += 1
            else:
                assert abs(flux) < 1e-12
    assert hot == 1


def test_gauge_cut_direction_is_invisible(orc):
    d, lo, hi = plane(10.0, 64, 10.0, 64)
    phi, xs, ys = 2.4, 0.31, -0.17
    a, _ = packet(d, lo, hi, -3.0, 0.5, 2.0, 1.0)
    b = a.copy().reshape(64, 64)
    x = lo[0] + np.arange(64) * 20.0 / 64
    isol = max(i for i in range(63) if x[i] <= xs < x[i + 1])
    b[:, isol + 1:] *= complex(math.cos(phi), -math.sin(phi))
    ra, _ = orc.lat2d_evolve(np.zeros(64 * 64), d, lo, hi, 0.02, a, 400, links=2, a=phi,
                             xs=xs, ys=ys, cut_up=True)
    rb, _ = orc.lat2d_evolve(np.zeros(64 * 64), d, lo, hi, 0.02, b.reshape(a.shape), 400,
                             links=2, a=phi, xs=xs, ys=ys, cut_up=False)
    assert np.abs(np.abs(ra) ** 2 - np.abs(rb) ** 2).max() < 1e-10


def test_uniform_field_fills_every_plaquette_equally(orc):
    d, lo, hi = plane(8.0, 32, 8.0, 32)
    b = 0.7
    lx = orc.lat2d_links(d, lo, hi, 1, b)
    want = b * 0.5 * 0.5
    for j in range(31):
        for i in range(31):
            assert abs(math.remainder(cmath.phase(_loop(lx, d, i, j)) - want, 2 * math.pi)) < 1e-12


def test_relax_finds_the_fock_darwin_ground(orc):
    d, lo, hi = plane(20.0, 128, 20.0, 128)
    w0, b = 0.5, 0.6
    x = lo[0] + np.arange(128) * 40.0 / 128
    yy, xx = np.meshgrid(x, x, indexing="ij")
    v = (0.5 * w0 * w0 * (xx ** 2 + yy ** 2)).ravel()
    psi, dv = packet(d, lo, hi, 1.0, -2.0, 3.0, 0.0)
    out, e = orc.lat2d_evolve(v, d, lo, hi, 0.02, psi, 3000, relax=True, links=1, a=b)
    omega = math.sqrt(w0 * w0 + 0.25 * b * b)
    assert abs(e - omega) < 0.03 * omega
    assert abs((np.abs(out) ** 2).sum() * dv - 1.0) < 1e-9


@needs_ref
def test_lattice_bitwise_equals_reference(orc):
    d, lo, hi = (32, 24, 1), (-8.0, -6.0, -1.0), (8.0, 6.0, 1.0)
    v = orc.harmonic_3d(d, lo, hi, 0.5, (0, 0, 0))
    g = orc.gaussian_3d(d, lo, hi, (1.0, 0.5, 0), (1.5, 1.5, 1.0), (0.7, 0, 0)).reshape(d[::-1])
    for kw in [{}, dict(links=1, a=0.6), dict(links=2, a=1.9, xs=0.3, ys=-0.2),
               dict(links=2, a=1.9, xs=0.3, ys=-0.2, cut_up=False),
               dict(relax=True, links=1, a=0.6)]:
        mine, e = orc.lat2d_evolve(v, d, lo, hi, 0.02, g, 30, **kw)
        theirs, e_ref, lx = orc.Ref.lat2d_evolve(v, d, lo, hi, 0.02, g, 30, **kw)
        assert np.array_equal(mine, theirs) and e == e_ref
        assert np.array_equal(lx, orc.lat2d_links(d, lo, hi, kw.get("links", 0), kw.get("a", 0.0),
                                                  kw.get("xs", 0.0), kw.get("ys", 0.0),
                                                  kw.get("cut_up", True)))
