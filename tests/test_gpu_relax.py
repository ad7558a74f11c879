"""Imaginary-time relaxation on the GPU (plans created with
TDSE_FLAG_IMAGINARY_TIME) against the oracle's restatement of
ImaginaryTimePropagator1D/3D (core/src/imaginary_time.ixx), which
tests/test_oracle_relax.py pins to the reference bit for bit.

Tolerances (max |psi_gpu - psi_oracle| / max |psi_oracle|):
  FP64 state:  1e-11
  FP32 state:  2e-5
The energy estimate -ln(norm)/(2 dtau) of the last step is compared with the
oracle's pre-renormalisation norm of the same step (relative 1e-10 FP64,
1e-5 FP32), and the relaxed states with the closed forms of the reference's
own tests (tests/imaginary_time_test.cpp, tests/deflation_test.cpp).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-11, "fp32": 2e-5}
ETOL = {"fp64": 1e-10, "fp32": 1e-5}


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def grid3(t, n, lo, hi):
    return t.Grid3D(t.Grid1D(lo[0], hi[0], n[0]), t.Grid1D(lo[1], hi[1], n[1]),
                    t.Grid1D(lo[2], hi[2], n[2]))


CASES_1D = [(512, 200), (1024, 100), (65536, 20), (96 * 512, 10)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n,steps", CASES_1D)
def test_relax_1d_matches_oracle(orc, tdse, n, steps, prec):
    lo, hi = (-16.0, 16.0) if n <= 1024 else (-100.0, 100.0)
    print(f"case relax 1d n={n} steps={steps} prec={prec}", flush=True)
    v = orc.soft_coulomb_1d(lo, hi, n, 1.0, 1.0, 0.0)
    psi0 = orc.gaussian_1d(lo, hi, n, 0.5, 1.5, 0.3)
    if n & (n - 1) == 0:
        want, last = orc.it1d_relax(v, lo, hi, 0.005, psi0, steps)
    else:
        want, last = orc.it3d_relax(v, (n, 1, 1), (lo, 0, 0), (hi, 1, 1), 0.005,
                                    psi0.reshape(1, 1, n), steps, any_size=True)
        want = want.ravel()
    p = tdse.Propagator(tdse.Grid1D(lo, hi, n), v, 0.005, precision=prec, imaginary_time=True)
    p.upload(psi0)
    p.step(steps)
    got = p.download()
    err = rel(got, want)
    st = p.relax_stats()
    e_want = -math.log(last) / (2 * 0.005)
    print(f"  rel err {err:.3e}; energy {st['energy']:.12f} vs {e_want:.12f}", flush=True)
    assert err < TOL[prec]
    assert abs(st["energy"] - e_want) < ETOL[prec] * max(1.0, abs(e_want))
    assert abs(st["norm_sq"] - last) < ETOL[prec] * last


CASES_3D = [
    ((32, 32, 32), (-8, -8, -8), (8, 8, 8), 40, 1.0),
    ((64, 32, 16), (-10, -6, -4), (10, 6, 4), 25, 0.7),
    ((128, 64, 1), (-16, -8, 0), (16, 8, 1), 30, 1.0),
    ((96, 32, 16), (-12, -6, -4), (12, 6, 4), 20, 1.0),
]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES_3D, ids=lambda c: "x".join(map(str, c[0])))
def test_relax_3d_matches_oracle(orc, tdse, case, prec):
    dims, lo, hi, steps, mass = case
    print(f"case relax 3d dims={dims} prec={prec}", flush=True)
    v = orc.soft_coulomb_3d(dims, lo, hi, 1.0, 0.8, (0.3, -0.2, 0.0))
    psi0 = orc.gaussian_3d(dims, lo, hi, (1.0, 0.5, 0.0), (1.4, 1.2, 1.0),
                           (0.3, 0.0, 0.2)).reshape(dims[::-1])
    pow2 = all(d & (d - 1) == 0 for d in dims)
    want, last = orc.it3d_relax(v, dims, lo, hi, 0.02, psi0, steps, mass=mass,
                                any_size=not pow2)
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, 0.02, mass=mass, precision=prec,
                        imaginary_time=True)
    p.upload(psi0)
    p.step(steps)
    err = rel(p.download(), want)
    st = p.relax_stats()
    e_want = -math.log(last) / (2 * 0.02)
    print(f"  rel err {err:.3e}; energy {st['energy']:.12f} vs {e_want:.12f}", flush=True)
    assert err < TOL[prec]
    assert abs(st["energy"] - e_want) < ETOL[prec] * max(1.0, abs(e_want))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_relax_split_calls_match_one_call(orc, tdse, prec):
    dims, lo, hi = (32, 16, 16), (-8, -4, -4), (8, 4, 4)
    v = orc.harmonic_3d(dims, lo, hi, 0.8, (0, 0, 0))
    psi0 = orc.gaussian_3d(dims, lo, hi, (0.7, 0, 0), (1.5, 1.5, 1.5), (0, 0, 0))
    g = grid3(tdse, dims, lo, hi)
    a = tdse.Propagator(g, v, 0.02, precision=prec, imaginary_time=True)
    b = tdse.Propagator(g, v, 0.02, precision=prec, imaginary_time=True)
    a.upload(psi0)
    b.upload(psi0)
    a.step(30)
    for k in (1, 2, 7, 20):
        b.step(k)
    assert rel(b.download(), a.download()) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_deflated_relax_matches_oracle(orc, tdse, prec):
    dims, lo, hi = (32, 32, 32), (-8.0,) * 3, (8.0,) * 3
    v = orc.harmonic_3d(dims, lo, hi, 1.0, (0.0, 0.0, 0.0))
    g0 = orc.gaussian_3d(dims, lo, hi, (0, 0, 0), (1.2, 1.2, 1.2), (0, 0, 0)).reshape(dims[::-1])
    ground, _ = orc.it3d_relax(v, dims, lo, hi, 0.05, g0, 200)
    guess = orc.gaussian_3d(dims, lo, hi, (1.0, 0.5, 0.0), (1.2, 1.2, 1.2),
                            (0, 0, 0)).reshape(dims[::-1])
    want, last = orc.it3d_relax(v, dims, lo, hi, 0.05, guess, 60, lower=[ground])
    g = grid3(tdse, dims, lo, hi)
    low = tdse.Propagator(g, None, 0.05, precision=prec)
    low.upload(ground)
    p = tdse.Propagator(g, v, 0.05, precision=prec, imaginary_time=True)
    p.upload(guess)
    print(f"case deflated 32^3 prec={prec}", flush=True)
    p.relax_deflated([low], 60)
    got = p.download()
    err = rel(got, want)
    print(f"  rel err {err:.3e}", flush=True)
    assert err < (TOL[prec] if prec == "fp64" else 1e-4)
    assert abs(p.relax_stats()["norm_sq"] - last) < (1e-9 if prec == "fp64" else 1e-5) * last


def test_ground_and_excited_states_closed_form(orc, tdse):
    """tests/deflation_test.cpp on the GPU: E0 = 3/2, E1 = 5/2 (2e-2), the
    excited state orthogonal to the ground state (1e-6) and normalised (1e-12)."""
    dims, lo, hi = (32, 32, 32), (-8.0,) * 3, (8.0,) * 3
    g = grid3(tdse, dims, lo, hi)
    v = tdse.harmonic_potential(g, 1.0, (0.0, 0.0, 0.0))
    relax = tdse.ImaginaryTimePropagator3D(g, v, 0.05)
    ground = tdse.gaussian_wavepacket(g, (0, 0, 0), (1.2, 1.2, 1.2), (0, 0, 0))
    relax.relax(ground, 600)
    guess = tdse.gaussian_wavepacket(g, (1.0, 0.5, 0.0), (1.2, 1.2, 1.2), (0, 0, 0))
    excited = guess.copy()
    relax.relax_deflated(excited, [ground], 600)
    e0 = orc.mean_energy_3d(ground, dims, lo, hi, v)
    e1 = orc.mean_energy_3d(excited, dims, lo, hi, v)
    print(f"E0 {e0:.6f} E1 {e1:.6f}", flush=True)
    assert abs(e0 - 1.5) < 2e-2 and abs(e1 - 2.5) < 2e-2
    ov = orc.inner_product(ground, excited, 0.5 ** 3)
    assert abs(ov.real) < 1e-6 and abs(ov.imag) < 1e-6
    assert abs(orc.norm_sq(excited, 0.5 ** 3) - 1.0) < 1e-12


def test_harmonic_ground_state_1d_closed_form(orc, tdse):
    """tests/imaginary_time_test.cpp::FindsHarmonicGroundState on the GPU."""
    lo, hi, n = -16.0, 16.0, 512
    g = tdse.Grid1D(lo, hi, n)
    v = tdse.harmonic_potential(g, 1.0, 0.0)
    psi = tdse.gaussian_wavepacket(g, 1.0, 2.0, 0.0)
    r = tdse.ImaginaryTimePropagator1D(g, v, 0.005)
    r.relax(psi, 6000)
    assert abs(orc.mean_energy_1d(psi, lo, hi, v) - 0.5) < 1e-5
    x = lo + np.arange(n) * (hi - lo) / n
    w = np.abs(psi) ** 2
    m = (w * x).sum() / w.sum()
    assert abs(math.sqrt((w * x * x).sum() / w.sum() - m * m) - 1 / math.sqrt(2)) < 1e-5
    assert abs(orc.mean_position_1d(psi, lo, hi)) < 1e-8
    assert abs(r.stats()["energy"] - 0.5) < 1e-4


def test_imaginary_plans_reject_mask_and_batch(orc, tdse):
    dims, lo, hi = (16, 16, 16), (-4.0,) * 3, (4.0,) * 3
    g = grid3(tdse, dims, lo, hi)
    v = orc.harmonic_3d(dims, lo, hi, 1.0, (0, 0, 0))
    p = tdse.Propagator(g, v, 0.05, imaginary_time=True)
    with pytest.raises(Exception):
        p.set_mask(orc.absorbing_mask_3d(dims, lo, hi, 1.0))
    with pytest.raises(Exception):
        tdse.Batch([p])
    q = tdse.Propagator(g, v, 0.05)
    with pytest.raises(ValueError):
        q.relax_stats()
