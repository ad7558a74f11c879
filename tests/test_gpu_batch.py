"""Batched uneven instances: one tdse_batch over plans of different shapes
must give, for every plan, exactly what stepping that plan alone gives (the
batched kernels run the same per-tile code), with and without hipGraphs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make(tdse, s):
    p = tdse.Propagator(s.grid(), s.build_potential(), s.dt, precision=s.precision,
                        mask=s.build_mask())
    p.upload(s.build_psi0())
    return p


@pytest.mark.parametrize("use_graph", [
    False,
    pytest.param(True, marks=pytest.mark.skip(
        reason="captured-graph replay crashed the host process in tdse_batch_step; "
               "switched off (TDSE_E_UNSUPPORTED) until a cause is shown")),
])
def test_batch_equals_solo(tdse, use_graph):
    from shoeridingerequationsolver_amd import workloads
    shapes = [(32, 32, 32), (96, 16, 8), (64, 64, 1), (128, 32, 1), (65536, 1, 1), (1024, 1, 1),
              (256, 1, 1), (16, 32, 64)]
    solves = workloads.batch_instances(seed=11, count=12, steps=23, shapes=shapes)
    solves += workloads.batch_instances(seed=12, count=4, steps=23, shapes=shapes,
                                        precision="fp64")
    solo = [make(tdse, s) for s in solves]
    batched = [make(tdse, s) for s in solves]
    for p in solo:
        p.step(23)
    b = tdse.Batch(batched)
    print(f"{len(solves)} plans, {b.launches_per_round()} batched launches per round", flush=True)
    b.step(23, use_graph=use_graph)
    for s, p, q in zip(solves, solo, batched):
        a, c = p.download(), q.download()
        assert np.array_equal(a, c), f"{s.name} {s.n} {s.precision}"


def test_batch_split_calls_and_reuse(tdse):
    """step(a); step(b) on a batch == step(a + b) per plan to rounding."""
    from shoeridingerequationsolver_amd import workloads
    solves = workloads.batch_instances(seed=13, count=6, steps=10,
                                       shapes=[(32, 32, 16), (64, 64, 1), (4096, 1, 1)])
    one = [make(tdse, s) for s in solves]
    two = [make(tdse, s) for s in solves]
    b1, b2 = tdse.Batch(one), tdse.Batch(two)
    b1.step(17)
    b2.step(9)
    b2.step(8)
    for p, q in zip(one, two):
        a, c = p.download(), q.download()
        assert np.abs(a - c).max() <= 1e-5 * np.abs(a).max()


def test_batch_graph_replay_is_switched_off(tdse):
    from shoeridingerequationsolver_amd import workloads
    s = workloads.batch_instances(seed=15, count=2, steps=1, shapes=[(32, 32, 32)])
    b = tdse.Batch([make(tdse, x) for x in s])
    with pytest.raises(tdse.TdseError):
        b.step(3, use_graph=True)


def test_batch_rejects_bad_members(tdse):
    from shoeridingerequationsolver_amd import workloads
    s = workloads.batch_instances(seed=14, count=1, steps=1, shapes=[(32, 32, 32)])[0]
    p = make(tdse, s)
    with pytest.raises(ValueError):
        tdse.Batch([p, p])


def test_batch_runs_each_plans_variants(tdse):
    """Members swept in the kernel variants their plans hold (set here, as
    the autotuner would) give exactly the solo plans' states in the same
    variants; members of one shape in different variants share no launch."""
    from shoeridingerequationsolver_amd import workloads
    shapes = [(64, 64, 64), (128, 64, 1), (96, 64, 64), (65536, 1, 1),  # lengths with variants
              (192, 64, 64), (64, 192, 96)]
    solves = workloads.batch_instances(seed=17, count=12, steps=15, shapes=shapes)
    solo, members = [], []
    for i, s in enumerate(solves):
        for group in (solo, members):
            p = make(tdse, s)
            names = list(p.variants)
            p.set_variants({k: (i + j) % 4 for j, k in enumerate(names)})
            group.append(p)
    for p in solo:
        p.step(15)
    b = tdse.Batch(members)
    print(f"{len(members)} plans with variants, {b.launches_per_round()} launches per round",
          flush=True)
    b.step(15)
    for s, p, q in zip(solves, solo, members):
        assert p.variants == q.variants
        assert np.array_equal(p.download(), q.download()), f"{s.name} {s.n} {p.variants}"
