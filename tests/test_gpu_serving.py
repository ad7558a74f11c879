"""Continuous batching: requests served in a mix (admitted under an HBM
budget, stepped together in quanta, joining while others run) give the same
state as each request solved alone, within the precision's tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(65536, 1, 1), (256, 256, 1), (32, 32, 32), (64, 64, 64), (96, 96, 96), (1024, 1, 1)]


@pytest.mark.parametrize("policy", ["streams", "batch"])
def test_mix_equals_alone(tdse, policy):
    from shoeridingerequationsolver_amd import serving
    trace = serving.request_stream(seed=21, count=14, rate=400.0, max_steps=160, shapes=SHAPES)
    # a budget that holds only a few requests at once, so admission waits
    need = sorted(serving.plan_bytes(s.grid(), s.precision, s.mask_width > 0) for _, s in trace)
    budget = need[-1] + need[-2] + need[-3]
    srv = serving.Server(budget, quantum=40, policy=policy, chunk_ms=0.5)
    rep = srv.run(trace)
    srv.close()
    summ = rep.summary()
    print(summ, flush=True)
    assert summ["solves"] == len(trace)
    assert rep.quanta > 1 and rep.max_active > 1
    assert rep.peak_bytes <= budget
    by_name = {r.name: r for r in rep.results}
    for _, s in trace:
        p = tdse.Propagator(s.grid(), None, s.dt, precision=s.precision)
        p.build(s.recipe())
        p.step(s.steps)
        alone = p.download()
        mixed = by_name[s.name].psi
        err = np.abs(mixed - alone).max() / np.abs(alone).max()
        tol = 1e-12 if s.precision == "fp64" else 2e-6
        assert err < tol, f"{s.name} {s.n} {s.precision} steps={s.steps}: {err:.3e}"


def test_budget_too_small_is_rejected(tdse):
    from shoeridingerequationsolver_amd import serving
    trace = serving.request_stream(seed=22, count=2, rate=100.0, max_steps=10,
                                   shapes=[(64, 64, 64)])
    srv = serving.Server(1 << 20, quantum=10)
    with pytest.raises(ValueError):
        srv.run(trace)
    srv.close()


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 32, 1), (4096, 1, 1), (96, 8, 4)])
@pytest.mark.parametrize("kind", ["harmonic", "regularized_coulomb", "barrier"])
def test_device_build_matches_host_builders(tdse, shape, kind):
    """Inputs built on the GPU from a recipe equal the host builders (which
    equal the reference bit for bit) to a few ulps.  On the 96-point axis over
    [-8, 8) (spacing 1/6, inexact) the barrier's edges x = 0 and x = 2 are grid
    points whose coordinate a fused multiply-add would put on the other side."""
    from shoeridingerequationsolver_amd import workloads
    dims = sum(1 for v in shape if v > 1)
    if kind == "regularized_coulomb" and dims == 1:
        pytest.skip("no 1D Coulomb builder")
    half = 8.0
    lo = tuple(-half if v > 1 else 0.0 for v in shape)
    hi = tuple(half if v > 1 else 1.0 for v in shape)
    params = {"harmonic": {"omega": 0.4}, "regularized_coulomb": {"Z": 1.3},
              "barrier": {"V0": 1.1, "x_lo": 0.0, "x_hi": 2.0}}[kind]
    s = workloads.Solve("b", shape, lo, hi, kind, params,
                        tuple(0.7 if v > 1 else 0.0 for v in shape),
                        tuple(1.3 if v > 1 else 1.0 for v in shape),
                        tuple(0.4 if v > 1 else 0.0 for v in shape), 0.01, 10, "fp64",
                        mask_width=2.0 if kind == "barrier" else 0.0)
    a = tdse.Propagator(s.grid(), s.build_potential(), s.dt, precision="fp64", mask=s.build_mask())
    a.upload(s.build_psi0())
    b = tdse.Propagator(s.grid(), None, s.dt, precision="fp64")
    b.build(s.recipe())
    pa, pb = a.download(), b.download()
    assert np.abs(pa - pb).max() < 1e-13 * np.abs(pa).max() * 100
    a.step(10)
    b.step(10)
    assert np.abs(a.download() - b.download()).max() < 1e-12 * np.abs(pa).max()


def test_large_device_builds_are_complete(tdse):
    """Plan memory is cleared and built on the plan's own stream: a large
    FP64 plan built on the GPU equals the host builders everywhere, every
    time (a clear left on another stream could overwrite part of the build)."""
    from shoeridingerequationsolver_amd import workloads
    s = workloads.hydrogen(seed=7, n=256, steps=3, precision="fp64")
    s.potential, s.params = "regularized_coulomb", {"Z": 1.0}
    host_psi = s.build_psi0()
    a = tdse.Propagator(s.grid(), s.build_potential(), s.dt, precision="fp64")
    a.upload(host_psi)
    a.step(3)
    want = a.download()
    a.close()
    scale = np.abs(host_psi).max()
    for rep in range(3):
        print(f"case 256^3 fp64 device build rep {rep}", flush=True)
        b = tdse.Propagator(s.grid(), None, s.dt, precision="fp64")
        b.build(s.recipe())
        assert np.abs(b.download() - host_psi).max() < 1e-11 * scale
        b.step(3)
        assert np.abs(b.download() - want).max() < 1e-11 * scale
        b.close()


@pytest.mark.parametrize("shape", [(32, 32, 16), (64, 32, 1), (96, 96, 1), (4096, 1, 1)])
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("how", ["build", "set_mask"])
def test_plan_bytes_equals_held_bytes(tdse, shape, prec, how):
    """The admission estimate tdse_plan_bytes equals what a masked plan holds
    once its inputs are built (device recipe) or set (host mask), mask ramps
    included, so a server filling its budget to the byte never meets
    TDSE_E_OOM mid-serve."""
    from shoeridingerequationsolver_amd import serving, workloads
    ctx = tdse.Context(0)
    dims = sum(1 for v in shape if v > 1)
    half = 8.0
    lo = tuple(-half if v > 1 else 0.0 for v in shape)
    hi = tuple(half if v > 1 else 1.0 for v in shape)
    s = workloads.Solve("m", shape, lo, hi, "barrier", {"V0": 1.0, "x_lo": 0.0, "x_hi": 2.0},
                        tuple(0.0 for _ in shape), tuple(1.5 if v > 1 else 1.0 for v in shape),
                        tuple(0.0 for _ in shape), 0.01, 4, prec, mask_width=2.0)
    want = serving.plan_bytes(s.grid(), prec, with_mask=True)
    if how == "build":
        p = tdse.Propagator(s.grid(), None, s.dt, precision=prec, context=ctx)
        p.build(s.recipe())
    else:
        p = tdse.Propagator(s.grid(), s.build_potential(), s.dt, precision=prec,
                            mask=s.build_mask(), context=ctx)
    assert ctx.bytes_in_use() == want, f"{shape} {prec} {how} dims={dims}"
    p.step(4)
    p.close()
    assert ctx.bytes_in_use() == 0
    ctx.close()
