"""Bitwise repeatability of the GPU reductions: the reference's thread pool
combines fixed chunks in a fixed order, so a reduction gives the same bits on
every run (core/src/parallel.ixx:179-238; tests/parallel_test.cpp:54
BitwiseDeterministicAcrossRuns).  The GPU reductions here are fixed-order
trees with no atomics; these tests call them twice, across plans holding the
same state, across a solo plan and a batch member, and across kernel
variants that change only the block size of the kernel finishing the
imaginary-time norm."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALL = 1 | 2 | 4 | 8


def _grid(t, n, half):
    ax = [t.Grid1D(-half, half, v) if v > 1 else t.Grid1D(0.0, 1.0, 1) for v in n]
    return ax[0] if n[1] == n[2] == 1 else t.Grid3D(*ax)


def _case(t, n, prec):
    from shoeridingerequationsolver_amd import workloads
    dims = sum(1 for v in n if v > 1)
    half = {3: 10.0, 2: 20.0, 1: 60.0}[dims]
    s = workloads.Solve("d", n, tuple(-half if v > 1 else 0.0 for v in n),
                        tuple(half if v > 1 else 1.0 for v in n), "harmonic", {"omega": 0.3},
                        tuple(1.0 if v > 1 else 0.0 for v in n),
                        tuple(1.7 if v > 1 else 1.0 for v in n),
                        tuple(0.4 if v > 1 else 0.0 for v in n), 0.01, 30, prec, mask_width=2.0)
    return s


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("n", [(64, 32, 32), (256, 128, 1), (65536, 1, 1),
                               (512, 1, 1)])
def test_observables_repeat_bitwise(tdse, prec, n):
    s = _case(tdse, n, prec)
    a = tdse.Propagator(s.grid(), None, s.dt, precision=prec)
    b = tdse.Propagator(s.grid(), None, s.dt, precision=prec)
    for p in (a, b):
        p.build(s.recipe())
        p.step(s.steps)
    assert np.array_equal(a.download(), b.download())
    what = ALL | (16 if n[1] == n[2] == 1 else 0)
    first = a.observe(what)
    print(f"case {prec} {n}: norm {first['norm']:.17g} E {first['energy']:.17g}", flush=True)
    _same(first, a.observe(what))      # the same plan twice
    _same(first, b.observe(what))      # another plan holding the same state
    s1 = a.probability_in_slab(0, -1.0, 3.0)
    assert s1 == a.probability_in_slab(0, -1.0, 3.0) == b.probability_in_slab(0, -1.0, 3.0)


def test_batch_member_observables_equal_solo(tdse):
    from shoeridingerequationsolver_amd import workloads
    solves = workloads.batch_instances(seed=31, count=6, steps=12,
                                       shapes=[(32, 32, 32), (64, 64, 1), (65536, 1, 1)])
    solo = []
    for s in solves:
        p = tdse.Propagator(s.grid(), None, s.dt, precision=s.precision)
        p.build(s.recipe())
        p.step(s.steps)
        solo.append(p)
    members = []
    for s in solves:
        p = tdse.Propagator(s.grid(), None, s.dt, precision=s.precision)
        p.build(s.recipe())
        members.append(p)
    b = tdse.Batch(members)
    b.step(12)
    b.sync()
    for s, p, q in zip(solves, solo, members):
        print(f"case batch member {s.name} {s.n}", flush=True)
        _same(p.observe(ALL), q.observe(ALL))
    b.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_relax_stats_repeat_bitwise_across_block_sizes(tdse, prec):
    """Imaginary time: the per-step norm is finished by the sweep after the
    measuring one (block 0's preamble) or by its own kernel (a call's last
    step).  Kernel variant 2 of the y sweep doubles its block size and
    changes nothing else, so relaxing with it must give the same bits."""
    n = (64, 64, 64)
    grid = _grid(tdse, n, 8.0)
    v = tdse.harmonic_potential(grid, 0.8, (0.3, 0.0, 0.0))
    psi0 = tdse.gaussian_wavepacket(grid, (1.0, 0.5, 0.0), (1.5, 1.2, 1.0), (0.2, 0.0, 0.0))
    out = []
    for variants in ({"y_fwd": 0}, {"y_fwd": 2}, {"y_fwd": 0}):
        p = tdse.Propagator(grid, v, 0.01, precision=prec, imaginary_time=True)
        full = {k: 0 for k in p.variants}
        full.update(variants)
        p.set_variants(full)
        p.upload(psi0)
        p.step(25)
        p.step(7)
        out.append((p.download(), p.relax_stats(), p.observe(ALL)))
        p.close()
    for psi, st, ob in out[1:]:
        assert np.array_equal(psi, out[0][0])
        assert st == out[0][1]
        _same(ob, out[0][2])
    print(f"case relax {prec}: E {out[0][1]['energy']:.17g}", flush=True)
