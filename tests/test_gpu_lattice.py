"""The Peierls lattice on the GPU (an opening row sweep, then whole steps on
2D tiles, two per launch) against the oracle's restatement of
PeierlsLattice2D (core/src/lattice2d.ixx), pinned in
tests/test_oracle_lattice.py.  FP64 1e-11, FP32 2e-5 relative to max |psi|;
lattice energies 1e-11 (FP64) / 2e-6 (FP32)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-11, "fp32": 2e-5}


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def grid3(t, n, lo, hi):
    return t.Grid3D(t.Grid1D(lo[0], hi[0], n[0]), t.Grid1D(lo[1], hi[1], n[1]),
                    t.Grid1D(lo[2], hi[2], n[2]))


def packet(d, lo, hi, x0, y0, sigma, k0):
    x = lo[0] + np.arange(d[0]) * (hi[0] - lo[0]) / d[0]
    y = lo[1] + np.arange(d[1]) * (hi[1] - lo[1]) / d[1]
    yy, xx = np.meshgrid(y, x, indexing="ij")
    env = np.exp(-((xx - x0) ** 2 + (yy - y0) ** 2) / (4 * sigma * sigma))
    psi = (env * np.exp(1j * k0 * xx)).astype(np.complex128).reshape(1, d[1], d[0])
    return psi / np.abs(psi).max()


CASES = [
    ((64, 64, 1), (-10.0, -10.0, -1.0), (10.0, 10.0, 1.0), {}, 40),
    ((100, 70, 1), (-12.0, -9.0, -1.0), (12.0, 9.0, 1.0), dict(links=1, a=0.5), 30),
    ((128, 130, 1), (-16.0, -16.0, -1.0), (16.0, 16.0, 1.0),
     dict(links=2, a=1.9, xs=0.31, ys=-0.17), 25),
    ((96, 59, 1), (-12.0, -8.0, -1.0), (12.0, 8.0, 1.0),
     dict(links=2, a=2.4, xs=0.31, ys=-0.17, cut_up=False), 20),
    ((1024, 40, 1), (-60.0, -8.0, -1.0), (60.0, 8.0, 1.0), dict(links=1, a=0.2), 10),
    # the widest row (2048), odd and tiny widths, fewer rows than one step
    # tile, one row past a tile, and solenoid cuts across tile edges
    ((2048, 20, 1), (-60.0, -6.0, -1.0), (60.0, 6.0, 1.0), dict(links=1, a=0.1), 6),
    ((257, 33, 1), (-16.0, -8.0, -1.0), (16.0, 8.0, 1.0),
     dict(links=2, a=1.3, xs=0.9, ys=-1.1), 12),
    ((6, 7, 1), (-3.0, -3.0, -1.0), (3.0, 3.0, 1.0), dict(links=1, a=0.7), 15),
    ((130, 17, 1), (-12.0, -6.0, -1.0), (12.0, 6.0, 1.0),
     dict(links=2, a=2.0, xs=-0.4, ys=0.2, cut_up=False), 9),
    # 512 step tiles and more (the FP32 tiles then take 512 threads): a
    # solenoid cut through tile edges, an odd step count (a one-step launch
    # last), and a uniform field
    ((768, 704, 1), (-30.0, -28.0, -1.0), (30.0, 28.0, 1.0),
     dict(links=2, a=1.7, xs=0.9, ys=-1.1), 5),
    ((1024, 512, 1), (-40.0, -20.0, -1.0), (40.0, 20.0, 1.0), dict(links=1, a=0.15), 8),
    # a strong field: x-link angles B hx y up to ~1.2e3 rad, reduced in the kernels
    ((96, 80, 1), (-30.0, -40.0, -1.0), (30.0, 40.0, 1.0), dict(links=1, a=48.0), 10),
]


def _set_links(p, kw):
    if kw.get("links") == 1:
        p.set_uniform_field(kw["a"])
    elif kw.get("links") == 2:
        p.set_solenoid(kw["a"], kw["xs"], kw["ys"], kw.get("cut_up", True))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"-{c[3].get('links', 0)}")
def test_lattice_step_matches_oracle(orc, tdse, case, prec):
    dims, lo, hi, kw, steps = case
    print(f"case lattice dims={dims} links={kw} prec={prec}", flush=True)
    v = orc.harmonic_3d(dims, lo, hi, 0.3, (0.5, 0, 0))
    psi0 = packet(dims, lo, hi, -2.0, 1.0, 2.0, 1.2)
    want, e_want = orc.lat2d_evolve(v, dims, lo, hi, 0.02, psi0, steps, **kw)
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, 0.02, precision=prec, lattice=True)
    _set_links(p, kw)
    p.upload(psi0)
    p.step(steps)
    got = p.download()
    err = rel(got, want)
    e = p.observe(tdse.capi.TDSE_OBS_NORM | tdse.capi.TDSE_OBS_ENERGY)["energy"]
    print(f"  rel err {err:.3e}; energy {e:.14f} vs {e_want:.14f}", flush=True)
    assert err < TOL[prec]
    assert abs(e - e_want) < (1e-11 if prec == "fp64" else 2e-6) * max(1.0, abs(e_want))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lattice_split_calls(orc, tdse, prec):
    """Calls of 1, 2 and 7 steps (each opens with its own half step) equal
    one run of 10 steps."""
    dims, lo, hi = (160, 96, 1), (-16.0, -10.0, -1.0), (16.0, 10.0, 1.0)
    kw = dict(links=1, a=0.4)
    v = orc.harmonic_3d(dims, lo, hi, 0.3, (0.5, 0, 0))
    psi0 = packet(dims, lo, hi, -2.0, 1.0, 2.0, 1.2)
    want, _ = orc.lat2d_evolve(v, dims, lo, hi, 0.02, psi0, 10, **kw)
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, 0.02, precision=prec, lattice=True)
    _set_links(p, kw)
    p.upload(psi0)
    for k in (1, 2, 7):
        p.step(k)
    assert rel(p.download(), want) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lattice_relax_matches_oracle(orc, tdse, prec):
    dims, lo, hi = (80, 64, 1), (-10.0, -8.0, -1.0), (10.0, 8.0, 1.0)
    v = orc.harmonic_3d(dims, lo, hi, 0.5, (0, 0, 0))
    psi0 = packet(dims, lo, hi, 1.0, -1.0, 2.0, 0.0)
    want, _ = orc.lat2d_evolve(v, dims, lo, hi, 0.02, psi0, 40, relax=True, links=1, a=0.6)
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, 0.02, precision=prec, lattice=True,
                        imaginary_time=True)
    p.set_uniform_field(0.6)
    p.upload(psi0)
    p.step(40)
    assert rel(p.download(), want) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lattice_relax_with_solenoid_matches_oracle(orc, tdse, prec):
    """Imaginary time with a solenoid's branch cut (one x link per cut row),
    on an odd-sized lattice of several step tiles."""
    dims, lo, hi = (90, 70, 1), (-12.0, -9.0, -1.0), (12.0, 9.0, 1.0)
    kw = dict(links=2, a=1.1, xs=0.4, ys=-0.6, cut_up=False)
    v = orc.harmonic_3d(dims, lo, hi, 0.4, (0, 0, 0))
    psi0 = packet(dims, lo, hi, 1.0, 0.5, 2.0, 0.3)
    want, _ = orc.lat2d_evolve(v, dims, lo, hi, 0.02, psi0, 25, relax=True, **kw)
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, 0.02, precision=prec, lattice=True,
                        imaginary_time=True)
    _set_links(p, kw)
    p.upload(psi0)
    p.step(25)
    assert rel(p.download(), want) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_lattice_with_mask_matches_oracle(orc, tdse, prec):
    dims, lo, hi = (64, 48, 1), (-10.0, -8.0, -1.0), (10.0, 8.0, 1.0)
    v = np.zeros(64 * 48)
    m = orc.absorbing_mask_3d(dims, lo, hi, 3.0)
    psi0 = packet(dims, lo, hi, 4.0, 0.0, 1.5, 2.0)
    want = psi0.ravel().copy()
    for _ in range(30):
        want, _ = orc.lat2d_evolve(v, dims, lo, hi, 0.02, want, 1)
        want = want * m
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, 0.02, precision=prec, mask=m, lattice=True)
    p.upload(psi0)
    p.step(30)
    assert rel(p.download().ravel(), want) < TOL[prec]


def test_fock_darwin_ground_on_gpu(orc, tdse):
    """tests/lattice2d_test.cpp::RelaxFindsTheFockDarwinGround through the facade."""
    dims, lo, hi = (128, 128, 1), (-20.0, -20.0, -1.0), (20.0, 20.0, 1.0)
    g = grid3(tdse, dims, lo, hi)
    w0, b = 0.5, 0.6
    x = lo[0] + np.arange(128) * 40.0 / 128
    yy, xx = np.meshgrid(x, x, indexing="ij")
    v = (0.5 * w0 * w0 * (xx ** 2 + yy ** 2)).ravel()
    lat = tdse.PeierlsLattice2D(g, v, 0.02)
    lat.set_uniform_field(b)
    psi = packet(dims, lo, hi, 1.0, -2.0, 3.0, 0.0)
    dv = (40.0 / 128) ** 2 * 2.0
    psi = psi / math.sqrt((np.abs(psi) ** 2).sum() * dv)
    lat.relax(psi, 3000)
    omega = math.sqrt(w0 * w0 + 0.25 * b * b)
    assert abs(lat.energy(psi) - omega) < 0.03 * omega
    assert abs((np.abs(psi) ** 2).sum() * dv - 1.0) < 1e-9


def test_lattice_rejections(tdse):
    g = grid3(tdse, (32, 32, 1), (-4, -4, -1), (4, 4, 1))
    p = tdse.Propagator(g, np.zeros(1024), 0.02, lattice=True)
    with pytest.raises(ValueError):
        p.set_drive((1, 0, 0), 0.1, 1.0)
    with pytest.raises(ValueError):  # link angles B hx y of 1e5 rad and beyond
        p.set_uniform_field(1e5)
    with pytest.raises(Exception):
        tdse.Batch([p])
    with pytest.raises(ValueError):
        tdse.Propagator(tdse.Grid3D(tdse.Grid1D(-1, 1, 8), tdse.Grid1D(-1, 1, 8),
                                    tdse.Grid1D(-1, 1, 2)), None, 0.02, lattice=True)
