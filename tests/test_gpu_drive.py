"""The dipole drive fused into the kick sweeps, through the C ABI, against the
oracle's restatement of ses::driven_step (core/src/drive.ixx:52-66), pinned in
tests/test_oracle_drive.py.  Tolerances as the parity tests: FP64 1e-11,
FP32 2e-5 relative to max |psi|."""
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


CASES = [
    ((32, 32, 32), (-8, -8, -8), (8, 8, 8), (0.3, 0.2, 1.0), 0.4, 0.9, 12),
    ((64, 32, 16), (-10, -6, -4), (10, 6, 4), (1.0, -0.5, 0.25), 0.3, 1.3, 10),
    ((128, 64, 1), (-16, -8, 0), (16, 8, 1), (0.6, 0.8, 0.0), 0.2, 0.5, 15),
    ((96, 32, 16), (-12, -6, -4), (12, 6, 4), (0.0, 0.0, 1.0), 0.5, 0.0, 8),
    ((1024, 1, 1), (-40, 0, 0), (40, 1, 1), (1.0, 0.0, 0.0), 0.05, 0.3, 40),
    ((65536, 1, 1), (-100, 0, 0), (100, 1, 1), (1.0, 0.0, 0.0), 0.02, 0.2, 6),
]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_driven_step_matches_oracle(orc, tdse, case, prec):
    dims, lo, hi, axis, amp, omega, steps = case
    print(f"case drive dims={dims} prec={prec}", flush=True)
    v = orc.harmonic_3d(dims, lo, hi, 0.3, (0.2, 0.0, 0.0))
    psi0 = orc.gaussian_3d(dims, lo, hi, (1.0, 0.0, 0.0), (1.3, 1.1, 0.9),
                           (0.4, 0.0, 0.0)).reshape(dims[::-1])
    pow2 = all(d & (d - 1) == 0 for d in dims)
    t0 = 0.37
    want = orc.driven_evolve(v, dims, lo, hi, 0.02, psi0, steps, axis, amp, omega, t0=t0,
                             any_size=not pow2)
    g = tdse.Grid1D(lo[0], hi[0], dims[0]) if dims[1] == dims[2] == 1 else grid3(tdse, dims, lo, hi)
    p = tdse.Propagator(g, v, 0.02, precision=prec)
    p.set_drive(axis, amp, omega)
    p.time = t0
    p.upload(psi0.reshape(p.shape))
    p.step(steps)
    err = rel(p.download().ravel(), want.ravel())
    print(f"  rel err {err:.3e}", flush=True)
    assert err < TOL[prec]
    assert abs(p.time - (t0 + steps * 0.02)) < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_driven_split_calls_follow_the_clock(orc, tdse, prec):
    dims, lo, hi = (32, 16, 16), (-8, -4, -4), (8, 4, 4)
    v = orc.harmonic_3d(dims, lo, hi, 0.5, (0, 0, 0))
    psi0 = orc.gaussian_3d(dims, lo, hi, (0.5, 0, 0), (1.2, 1.2, 1.2), (0, 0, 0))
    want = orc.driven_evolve(v, dims, lo, hi, 0.02, psi0.reshape(dims[::-1]), 30,
                             (1, 0.5, 0), 0.3, 1.1)
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, 0.02, precision=prec)
    p.set_drive((1, 0.5, 0), 0.3, 1.1)
    p.upload(psi0)
    for k in (1, 4, 10, 15):
        p.step(k)
    assert rel(p.download(), want) < TOL[prec]


def test_zero_amplitude_is_the_static_step_bitwise(orc, tdse):
    dims, lo, hi = (32, 32, 16), (-8, -8, -4), (8, 8, 4)
    v = orc.harmonic_3d(dims, lo, hi, 1.0, (0, 0, 0))
    psi0 = orc.gaussian_3d(dims, lo, hi, (1.0, 0, 0), (1, 1, 1), (0, 0, 0))
    g = grid3(tdse, dims, lo, hi)
    a = tdse.Propagator(g, v, 0.02)
    b = tdse.Propagator(g, v, 0.02)
    b.set_drive((0, 0, 1), 0.0, 0.5)
    a.upload(psi0)
    b.upload(psi0)
    a.step(10)
    b.step(10)
    assert np.array_equal(a.download(), b.download())


def test_constant_field_ehrenfest_on_gpu(orc, tdse):
    """tests/drive_test.cpp::ConstantFieldObeysEhrenfest through the GPU path."""
    dims, lo, hi = (32, 32, 32), (-16.0,) * 3, (16.0,) * 3
    g = grid3(tdse, dims, lo, hi)
    dt, e0, steps = 0.04, 0.05, 50
    p = tdse.Propagator(g, np.zeros(32 ** 3), dt)
    p.set_drive((0, 0, 1), e0, 0.0)
    p.upload(tdse.gaussian_wavepacket(g, (0, 0, 0), (1.5, 1.5, 1.5), (0, 0, 0)))
    p.step(steps)
    o = p.observe(15)
    t = steps * dt
    assert abs(o["momentum"][2] - (-e0 * t)) < 1e-8
    assert abs(o["position"][2] - (-0.5 * e0 * t * t)) < 1e-6
    assert abs(o["norm"] - 1.0) < 1e-12


def test_driven_plans_do_not_batch(orc, tdse):
    dims, lo, hi = (16, 16, 16), (-4.0,) * 3, (4.0,) * 3
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), orc.harmonic_3d(dims, lo, hi, 1, (0, 0, 0)), 0.02)
    p.set_drive((1, 0, 0), 0.1, 1.0)
    with pytest.raises(Exception):
        tdse.Batch([p])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_driven_step_facade_matches_oracle(orc, tdse, prec):
    """ses::driven_step(psi, prop, d, t0, n) through the Python mirror: the
    oracle's driven steps from t0, and the propagator static again after."""
    dims, lo, hi = (32, 32, 16), (-8.0, -8.0, -4.0), (8.0, 8.0, 4.0)
    g = tdse.Grid3D(tdse.Grid1D(lo[0], hi[0], dims[0]), tdse.Grid1D(lo[1], hi[1], dims[1]),
                    tdse.Grid1D(lo[2], hi[2], dims[2]))
    v = orc.harmonic_3d(dims, lo, hi, 0.4, (0.0, 0.0, 0.0))
    psi0 = orc.gaussian_3d(dims, lo, hi, (1.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                           (0.0, 0.5, 0.0)).reshape(dims[::-1])
    d = tdse.DipoleDrive((1.0, 0.5, 0.0), 0.3, 0.7)
    print(f"case driven_step facade prec={prec}", flush=True)
    want = orc.driven_evolve(v, dims, lo, hi, 0.02, psi0, 9, d.axis, d.amplitude, d.omega,
                             t0=1.3)
    prop = tdse.SplitOperator3D(g, v, 0.02, precision=prec)
    psi = psi0.copy()
    tdse.driven_step(psi, prop, d, 1.3, 9)
    tol = 1e-11 if prec == "fp64" else 2e-5
    assert rel(psi, want) < tol
    psi = psi0.copy()
    prop.step(psi, 9)
    assert rel(psi, orc.so3d_evolve(v, dims, lo, hi, 0.02, psi0, 9)) < tol


def _radial_mask(dims, lo, hi, width):
    axes = [lo[a] + np.arange(dims[a]) * (hi[a] - lo[a]) / dims[a] for a in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    d = np.clip((min(hi) - r) / width, 0.0, 1.0)
    return (np.sin(0.5 * np.pi * d) ** 2).ravel()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["ramps", "radial"])
def test_driven_step_with_mask_matches_oracle(orc, tdse, kind, prec):
    """Drive and absorbing mask together: each step is driven_step, then the
    caller's mask (the oracle steps one at a time).  A product-of-ramps mask
    takes the x sweeps' ramp path, a radial one the per-cell array."""
    dims, lo, hi = (64, 32, 16), (-10.0, -6.0, -4.0), (10.0, 6.0, 4.0)
    axis, amp, omega, steps, dt, t0 = (0.8, 0.0, 0.6), 0.3, 1.1, 12, 0.02, 0.21
    print(f"case drive+mask {kind} prec={prec}", flush=True)
    v = orc.harmonic_3d(dims, lo, hi, 0.3, (0.2, 0.0, 0.0))
    m = orc.absorbing_mask_3d(dims, lo, hi, 2.5) if kind == "ramps" else _radial_mask(dims, lo, hi, 2.5)
    psi0 = orc.gaussian_3d(dims, lo, hi, (5.0, 0.0, 0.0), (1.3, 1.1, 0.9),
                           (1.4, 0.0, 0.0)).reshape(dims[::-1])
    want = psi0.copy()
    for s in range(steps):
        want = orc.driven_evolve(v, dims, lo, hi, dt, want, 1, axis, amp, omega, t0=t0 + s * dt)
        want = (want.ravel() * m).reshape(want.shape)
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, dt, precision=prec, mask=m)
    p.set_drive(axis, amp, omega)
    p.time = t0
    p.upload(psi0)
    p.step(5)
    p.step(steps - 5)
    err = rel(p.download(), want)
    print(f"  rel err {err:.3e}", flush=True)
    assert err < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_drive_near_the_angle_bound(orc, tdse, prec):
    """A drive whose kick angles reach 2e4 rad (a fifth of the 1e5 rad bound)
    still matches the oracle; the angle's own FP64 rounding (ulp 3.6e-12)
    enters every kick, so the FP64 bound here is 1e-10.  Beyond the
    bound set_drive refuses (TDSE_E_UNSUPPORTED -> ValueError)."""
    dims, lo, hi = (64, 32, 16), (-10, -6, -4), (10, 6, 4)
    axis, dt, steps = (1.0, 0.5, 0.25), 0.02, 6
    max_u = sum(abs(a) * max(abs(l), abs(h)) for a, l, h in zip(axis, lo, hi))
    amp = 2.0e4 / (dt * max_u)
    print(f"case drive near bound {prec}: amplitude {amp:.1f}", flush=True)
    v = orc.harmonic_3d(dims, lo, hi, 0.3, (0.2, 0.0, 0.0))
    psi0 = orc.gaussian_3d(dims, lo, hi, (1.0, 0.0, 0.0), (1.3, 1.1, 0.9),
                           (0.4, 0.0, 0.0)).reshape(dims[::-1])
    want = orc.driven_evolve(v, dims, lo, hi, dt, psi0, steps, axis, amp, 0.7, t0=0.1)
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, dt, precision=prec)
    p.set_drive(axis, amp, 0.7)
    p.time = 0.1
    p.upload(psi0)
    p.step(steps)
    err = rel(p.download().ravel(), want.ravel())
    print(f"  rel err {err:.3e}", flush=True)
    assert err < (1e-10 if prec == "fp64" else TOL[prec])
    with pytest.raises(ValueError):
        p.set_drive(axis, 1.01e5 / (dt * max_u), 0.7)
    with pytest.raises(ValueError):
        p.set_drive(axis, -3.0e5 / (dt * max_u), 0.7)
    p.set_drive(axis, 0.9e5 / (dt * max_u), 0.7)  # below the bound: accepted
    p.close()
