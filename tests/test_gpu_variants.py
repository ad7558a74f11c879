"""Kernel variants and the autotuner, through the C ABI, against the oracle.

Every compiled variant (alternative radix list, wider tiles) of every sweep
must give the oracle's result to the same tolerance as the default kernels:
FP64 1e-11 and FP32 2e-5 relative to max |psi|, as in the parity tests.  The
variants only change how a line's points are split over lanes and how many
lines share a block, so memory stays in natural order between sweeps and any
mix of variants over the slots of one schedule is valid.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11
FP32_TOL = 2e-5


def grid3(t, n, lo, hi):
    return t.Grid3D(t.Grid1D(lo[0], hi[0], n[0]), t.Grid1D(lo[1], hi[1], n[1]),
                    t.Grid1D(lo[2], hi[2], n[2]))


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def uniform_variant(prop, v):
    """Variant v on every slot that has it compiled, 0 elsewhere."""
    names = list(prop.variants)
    chosen = {k: 0 for k in names}
    for k in names:
        trial = dict(chosen)
        trial[k] = v
        try:
            prop.set_variants(trial)
            chosen = trial
        except ValueError:
            pass
    prop.set_variants(chosen)
    return chosen


CASES = [
    # dims, lo, hi, steps: tuned lengths in every role (x contiguous, y strided
    # forward/inverse, z strided kinetic, 2D kinetic, four-step 1D)
    ((64, 96, 128), (-10, -12, -14), (10, 12, 14), 6),
    ((96, 64, 256), (-12, -8, -16), (12, 8, 16), 4),
    ((192, 192, 16), (-12, -12, -8), (12, 12, 8), 4),
    ((32, 16, 192), (-8, -8, -12), (8, 8, 12), 4),
    ((256, 128, 1), (-30, -15, 0), (30, 15, 1), 6),
    ((1024, 512, 1), (-60, -40, 0), (60, 40, 1), 3),
    ((65536, 1, 1), (-100, 0, 0), (100, 1, 1), 5),
]


def _inputs(orc, dims, lo, hi):
    if dims[1] == 1 and dims[2] == 1:
        v = orc.harmonic_1d(lo[0], hi[0], dims[0], 0.1)
        m = orc.absorbing_mask_1d(lo[0], hi[0], dims[0], 8.0)
        psi0 = orc.gaussian_1d(lo[0], hi[0], dims[0], -15.0, 1.0, 2.0)
        return v, m, psi0
    v = orc.harmonic_3d(dims, lo, hi, 0.3, (0.2, -0.1, 0.0))
    m = orc.absorbing_mask_3d(dims, lo, hi, 3.0)
    psi0 = orc.gaussian_3d(dims, lo, hi, (0.5, -0.3, 0.0), (1.3, 1.1, 0.9),
                           (0.7, -0.4, 0.3)).reshape(dims[::-1])
    return v, m, psi0


def _want(orc, dims, lo, hi, v, m, psi0, steps):
    if dims[1] == 1 and dims[2] == 1:
        return orc.so1d_evolve(v, lo[0], hi[0], 0.02, psi0, steps, mask=m)
    pow2 = all(d & (d - 1) == 0 for d in dims)
    return orc.so3d_evolve(v, dims, lo, hi, 0.02, psi0, steps, mask=m, any_size=not pow2)


def _prop(tdse, dims, lo, hi, v, m, prec):
    g = tdse.Grid1D(lo[0], hi[0], dims[0]) if dims[1] == dims[2] == 1 else grid3(tdse, dims, lo, hi)
    return tdse.Propagator(g, v, 0.02, precision=prec, mask=m)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_every_variant_matches_oracle(orc, tdse, case, prec):
    dims, lo, hi, steps = case
    v, m, psi0 = _inputs(orc, dims, lo, hi)
    want = _want(orc, dims, lo, hi, v, m, psi0, steps)
    tol = FP64_TOL if prec == "fp64" else FP32_TOL
    p = _prop(tdse, dims, lo, hi, v, m, prec)
    for var in range(4):
        chosen = uniform_variant(p, var)
        print(f"case dims={dims} prec={prec} variants={chosen}", flush=True)
        p.upload(psi0)
        p.step(steps)
        err = rel(p.download(), want)
        print(f"  rel err {err:.3e}", flush=True)
        assert err < tol
    p.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mixed_variants_match_oracle(orc, tdse, prec):
    dims, lo, hi = (64, 96, 128), (-10, -12, -14), (10, 12, 14)
    v, m, psi0 = _inputs(orc, dims, lo, hi)
    want = _want(orc, dims, lo, hi, v, m, psi0, 6)
    p = _prop(tdse, dims, lo, hi, v, m, prec)
    mix = {"x_first": 1, "y_fwd": 3, "z_kin": 2, "y_inv": 1, "x_mid": 2, "x_last": 3}
    print(f"case mixed prec={prec} {mix}", flush=True)
    p.set_variants(mix)
    assert p.variants == mix
    p.upload(psi0)
    p.step(6)
    assert rel(p.download(), want) < (FP64_TOL if prec == "fp64" else FP32_TOL)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_autotune_keeps_state_and_parity(orc, tdse, tmp_path, prec):
    dims, lo, hi = (96, 64, 256), (-12, -8, -16), (12, 8, 16)
    v, m, psi0 = _inputs(orc, dims, lo, hi)
    want = _want(orc, dims, lo, hi, v, m, psi0, 4)
    p = _prop(tdse, dims, lo, hi, v, m, prec)
    p.upload(psi0)
    before = p.download()
    cache = str(tmp_path / "tune.json")
    print(f"case autotune prec={prec}", flush=True)
    got = p.autotune(reps=3, cache=cache)
    print(f"  chosen {got}", flush=True)
    assert set(got) == set(p.variants) and all(0 <= x < 4 for x in got.values())
    assert p.variants == got
    assert np.array_equal(p.download(), before)  # the state is not touched
    p.step(4)
    assert rel(p.download(), want) < (FP64_TOL if prec == "fp64" else FP32_TOL)
    # the cache restores the same choice on a new plan without timing
    q = _prop(tdse, dims, lo, hi, v, m, prec)
    assert q.autotune(cache=cache) == got
    assert q.variants == got


def test_bad_variants_rejected(orc, tdse):
    dims, lo, hi = (16, 16, 16), (-4, -4, -4), (4, 4, 4)
    v, m, _ = _inputs(orc, dims, lo, hi)
    p = _prop(tdse, dims, lo, hi, v, m, "fp32")
    names = list(p.variants)
    with pytest.raises(ValueError):  # length 16 has no variants
        p.set_variants({k: 1 for k in names})
    with pytest.raises(ValueError):
        p.set_variants({k: 7 for k in names})
    with pytest.raises(ValueError):
        p.set_variants({names[0]: 0})
    assert p.variants == {k: 0 for k in names}
