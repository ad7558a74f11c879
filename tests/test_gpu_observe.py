"""On-device observables against the oracle's restatement of
core/src/observables.ixx (norm, position, width, momentum, energy), which
tests/test_oracle.py and tests/test_oracle_relax.py pin to the reference.

The k-space terms come from one read-only line-FFT sweep per axis (the 3D
ratio sum k_a^2 |psi_k|^2 / sum |psi_k|^2 is the same over the 1D spectra of
the lines along axis a); four-step 1D grids transform a scratch copy.
Tolerances: FP64 states 1e-11 on widths and moments (absolute, unit-scale
quantities) and 1e-10 on energies; FP32 states 2e-6 relative to the scale.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def grid3(t, n, lo, hi):
    return t.Grid3D(t.Grid1D(lo[0], hi[0], n[0]), t.Grid1D(lo[1], hi[1], n[1]),
                    t.Grid1D(lo[2], hi[2], n[2]))


CASES = [
    ((32, 16, 16), (-8, -4, -4), (8, 4, 4)),
    ((64, 64, 64), (-10, -10, -10), (10, 10, 10)),
    ((256, 128, 1), (-20, -10, 0), (20, 10, 1)),
    ((96, 32, 16), (-12, -6, -4), (12, 6, 4)),
    ((160, 192, 1), (-16, -16, 0), (16, 16, 1)),
]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_observables_3d_match_oracle(orc, tdse, case, prec):
    dims, lo, hi = case
    print(f"case observe dims={dims} prec={prec}", flush=True)
    v = orc.harmonic_3d(dims, lo, hi, 0.7, (0.5, 0, 0))
    psi = orc.gaussian_3d(dims, lo, hi, (1.0, 0.5, 0.0), (1.2, 1.1, 0.9),
                          (0.6, -0.4, 0.2)).reshape(dims[::-1])
    p = tdse.Propagator(grid3(tdse, dims, lo, hi), v, 0.02, precision=prec)
    p.upload(psi)
    if prec == "fp32":
        psi = p.download()  # the state as stored
    o = p.observe(tdse.capi.TDSE_OBS_NORM | tdse.capi.TDSE_OBS_POSITION |
                  tdse.capi.TDSE_OBS_ENERGY | tdse.capi.TDSE_OBS_MOMENTUM)
    tol = 1e-11 if prec == "fp64" else 2e-6
    cell = np.prod([(hi[a] - lo[a]) / dims[a] for a in range(3)])
    assert abs(o["norm"] - orc.norm_sq(psi, cell)) < tol
    np.testing.assert_allclose(o["position"], orc.mean_position_3d(psi, dims, lo, hi), atol=tol)
    np.testing.assert_allclose(o["sigma"], orc.sigma_position_3d(psi, dims, lo, hi), atol=tol)
    np.testing.assert_allclose(o["momentum"], orc.mean_momentum_3d(psi, dims, lo, hi), atol=tol)
    e = orc.mean_energy_3d(psi, dims, lo, hi, v)
    print(f"  energy {o['energy']:.14f} vs {e:.14f}", flush=True)
    assert abs(o["energy"] - e) < (1e-10 if prec == "fp64" else 2e-6 * max(1.0, abs(e)))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_box_probability_matches_oracle(orc, tdse, case, prec):
    """Box region probability against the oracle's restatement (pinned by the
    slab, the separable product and the Gaussian closed form) on a state that
    a few steps in a Coulomb well made non-separable.  Both sum the stored
    state's |psi|^2 in FP64, so FP32 states meet the same 1e-12.  One
    restricted axis gives the slab bit for bit; an axis of extent 1 is not
    restricted."""
    from shoeridingerequationsolver_amd import workloads
    dims, lo, hi = case
    print(f"case box dims={dims} prec={prec}", flush=True)
    one = [n == 1 for n in dims]
    s = workloads.Solve("box", dims, lo, hi, "regularized_coulomb", {"Z": 1.0},
                        tuple(0.0 if o else c for o, c in zip(one, (0.5, -0.3, 0.2))),
                        tuple(1.0 if o else c for o, c in zip(one, (1.3, 1.1, 0.9))),
                        tuple(0.0 if o else c for o, c in zip(one, (0.4, 0.2, -0.3))),
                        0.01, 5, prec)
    p = tdse.Propagator(s.grid(), None, s.dt, precision=prec)
    p.build(s.recipe())
    p.step(5)
    psi = p.download()
    a, b = (-1.0, -2.0, -1.5), (2.5, 1.0, 3.0)
    got = p.probability_in_box(a, b)
    want = orc.probability_in_box_3d(psi, dims, lo, hi, a, b)
    print(f"  box {got:.15f} vs {want:.15f}", flush=True)
    assert 0.0 < got < 1.0
    assert abs(got - want) < 1e-12
    inf = float("inf")
    assert p.probability_in_box((a[0], -inf, -inf), (b[0], inf, inf)) == \
        p.probability_in_slab(0, a[0], b[0])
    if dims[2] == 1:
        assert p.probability_in_box((a[0], a[1], 5.0), (b[0], b[1], 6.0)) == got
    with pytest.raises(tdse.TdseError):
        p.probability_in_box((float("nan"), 0.0, 0.0), (1.0, 1.0, 1.0))
    p.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("dims", [(64, 64, 64), (96, 32, 16), (65536, 1, 1), (160, 192, 1)],
                         ids=lambda d: "x".join(map(str, d)))
def test_peak_density_matches_the_state(tdse, dims, prec):
    """TDSE_OBS_PEAK: max |psi|^2 over the cells, without the cell volume (the
    reference engine's norm_and_peak().peak, solver/src/vk_engine.ixx:1665-1691),
    against the downloaded state (each term may round once differently from
    numpy's, so within 4 ulps); a max, so a second call gives the same bits;
    without the flag the field stays 0."""
    from shoeridingerequationsolver_amd import workloads
    print(f"case peak dims={dims} prec={prec}", flush=True)
    one = [n == 1 for n in dims]
    half = 100.0 if dims[1] == 1 else 16.0
    lo = tuple(0.0 if o else -half for o in one)
    hi = tuple(1.0 if o else half for o in one)
    pot = "harmonic" if dims[1] == 1 else "regularized_coulomb"
    params = {"omega": 0.1} if pot == "harmonic" else {"Z": 1.0}
    s = workloads.Solve("peak", dims, lo, hi, pot, params,
                        tuple(0.0 if o else c for o, c in zip(one, (0.5, -0.3, 0.2))),
                        tuple(1.0 if o else c for o, c in zip(one, (1.3, 1.1, 0.9))),
                        tuple(0.0 if o else c for o, c in zip(one, (0.4, 0.2, -0.3))),
                        0.01, 5, prec)
    p = tdse.Propagator(s.grid(), None, s.dt, precision=prec)
    p.build(s.recipe())
    p.step(5)
    psi = p.download()
    want = float((psi.real ** 2 + psi.imag ** 2).max())
    o = p.observe(tdse.capi.TDSE_OBS_NORM | tdse.capi.TDSE_OBS_PEAK)
    print(f"  peak {o['peak']:.17g} vs {want:.17g}", flush=True)
    assert abs(o["peak"] - want) <= 4 * np.finfo(np.float64).eps * want
    assert p.observe(tdse.capi.TDSE_OBS_PEAK)["peak"] == o["peak"]
    assert p.observe(tdse.capi.TDSE_OBS_NORM)["peak"] == 0.0
    p.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n,lo,hi", [(512, -16.0, 16.0), (1024, -40.0, 40.0),
                                     (65536, -100.0, 100.0)])
def test_observables_1d_match_oracle(orc, tdse, n, lo, hi, prec):
    print(f"case observe 1d n={n} prec={prec}", flush=True)
    v = orc.harmonic_1d(lo, hi, n, 0.1)
    psi = orc.gaussian_1d(lo, hi, n, 5.0, 2.0, 0.5)
    p = tdse.Propagator(tdse.Grid1D(lo, hi, n), v, 0.01, precision=prec)
    p.upload(psi)
    if prec == "fp32":
        psi = p.download()
    o = p.observe(15)
    tol = 1e-11 if prec == "fp64" else 2e-6
    assert abs(o["norm"] - orc.norm_sq(psi, (hi - lo) / n)) < tol
    assert abs(o["position"][0] - orc.mean_position_1d(psi, lo, hi)) < tol
    assert abs(o["sigma"][0] - orc.sigma_x_1d(psi, lo, hi)) < tol
    assert abs(o["momentum"][0] - orc.mean_momentum_1d(psi, lo, hi)) < tol
    assert abs(o["energy"] - orc.mean_energy_1d(psi, lo, hi, v)) < (1e-10 if prec == "fp64" else 2e-6)


def test_observables_of_relaxed_ground_state(orc, tdse):
    """Closed forms: the harmonic ground state has <k> = 0, width 1/sqrt(2 w)
    per axis and energy 3 w / 2 (to the grid and splitting accuracy)."""
    dims, lo, hi = (64, 64, 64), (-8.0,) * 3, (8.0,) * 3
    g = grid3(tdse, dims, lo, hi)
    v = tdse.harmonic_potential(g, 1.0, (0.0, 0.0, 0.0))
    p = tdse.Propagator(g, v, 0.01, imaginary_time=True)
    p.upload(tdse.gaussian_wavepacket(g, (0.5, 0.2, 0.0), (1.0, 0.9, 0.8), (0.3, 0, 0)))
    p.step(2000)
    o = p.observe(15)
    assert abs(o["energy"] - 1.5) < 1e-4
    np.testing.assert_allclose(o["sigma"], [2 ** -0.5] * 3, atol=1e-4)
    np.testing.assert_allclose(o["momentum"], [0, 0, 0], atol=1e-8)
    np.testing.assert_allclose(o["position"], [0, 0, 0], atol=1e-8)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n,xmin,xmax,omega", [(256, -20.0, 20.0, 0.25), (96, -15.0, 15.0, 0.5),
                                               (1024, -30.0, 30.0, 0.2),
                                               (65536, -100.0, 100.0, 4.0)])
def test_energy_variance_1d(tdse, orc, prec, n, xmin, xmax, omega):
    """TDSE_OBS_VARIANCE (core/src/observables.ixx:92-119) on the one-block
    line (256, 96, 1024) and the four-step line (65536): the oscillator's
    eigenstates give ~0, their equal mix (E1 - E0)^2 / 4 = w^2 / 4, a displaced
    packet the oracle's value (FP64 to 1e-11 relative, FP32 to 1e-5).  An FP32
    state carries rounding noise of relative size 2^-24 in every bin, on
    which H acts with up to 0.5 (pi/h)^2: FP32 results are checked within
    that floor, 4 (0.5 (pi/h)^2)^2 2^-48, on top."""
    from ho_states import ho_eigenstate
    grid = tdse.Grid1D(xmin, xmax, n)
    v = orc.harmonic_1d(xmin, xmax, n, omega)
    p = tdse.Propagator(grid, v, 0.01, precision=prec)
    states = {"ground": ho_eigenstate(xmin, xmax, n, omega, 0),
              "first": ho_eigenstate(xmin, xmax, n, omega, 1)}
    states["mix"] = states["ground"] + states["first"]
    states["packet"] = orc.gaussian_1d(xmin, xmax, n, 0.3 * xmax / 4, 1.2, 0.7)
    h = (xmax - xmin) / n
    floor = 0.0 if prec == "fp64" else 4.0 * (0.5 * (np.pi / h) ** 2) ** 2 * 2.0 ** -48
    for name, psi in states.items():
        p.upload(psi)
        got = p.observe(tdse.TDSE_OBS_NORM | tdse.TDSE_OBS_VARIANCE)["energy_variance"]
        want = orc.energy_variance_1d(psi, xmin, xmax, v)
        print(f"case {prec} n={n} {name}: Var(H) {got:.12e} oracle {want:.12e}", flush=True)
        if name in ("ground", "first"):
            assert abs(got) < (1e-10 if prec == "fp64" else 1e-9 * omega * omega) + floor
        elif name == "mix":
            assert abs(got - omega * omega / 4.0) < (1e-10 if prec == "fp64" else 1e-6) * omega ** 2 \
                + floor
        tol = 1e-11 if prec == "fp64" else 1e-5
        assert abs(got - want) <= tol * max(abs(want), omega * omega / 4.0) + floor
    p.close()


def test_energy_variance_high_level_eigenstate(tdse, orc):
    """The reference's lock (tests/observables_test.cpp:113-121): the level
    2500 eigenstate of w = 4 on the 65536 scene grid (E = 10002) has Var(H)
    at round-off, < 1e-8, through the four-step sweeps."""
    from ho_states import ho_eigenstate
    v = orc.harmonic_1d(-100.0, 100.0, 65536, 4.0)
    psi = ho_eigenstate(-100.0, 100.0, 65536, 4.0, 2500)
    p = tdse.Propagator(tdse.Grid1D(-100.0, 100.0, 65536), v, 0.001, precision="fp64")
    p.upload(psi)
    o = p.observe(tdse.TDSE_OBS_NORM | tdse.TDSE_OBS_ENERGY | tdse.TDSE_OBS_VARIANCE)
    print(f"case level 2500: Var(H) {o['energy_variance']:.3e}, <E> {o['energy']:.9f}", flush=True)
    assert o["energy_variance"] < 1e-8
    assert abs(o["energy"] - 10002.0) < 1e-2


def test_energy_variance_rejects_3d(tdse):
    g = tdse.Grid3D(tdse.Grid1D(-4, 4, 16), tdse.Grid1D(-4, 4, 16), tdse.Grid1D(-4, 4, 16))
    p = tdse.Propagator(g, np.zeros(16 ** 3), 0.01)
    with pytest.raises(tdse.TdseError):
        p.observe(tdse.TDSE_OBS_VARIANCE)
    p.close()
