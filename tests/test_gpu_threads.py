"""The boundary's threading contract: distinct plans may be driven from
distinct host threads.  Worker threads each create a plan in one shared
context, step it, observe it and destroy it while the main thread keeps
calling the context-wide wait (tdse_sync); every result equals the same solve
run alone on one thread, bit for bit (the same kernels on the plan's own
stream), and no call fails."""
import dataclasses
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(64, 64, 64), (96, 32, 16), (65536, 1, 1), (128, 128, 1), (1024, 1, 1), (32, 32, 32)]


def _solves():
    from shoeridingerequationsolver_amd import workloads
    base = workloads.batch_instances(seed=31, count=6, steps=10, shapes=SHAPES)
    return [dataclasses.replace(s, precision="fp64" if i % 2 else "fp32")
            for i, s in enumerate(base)]


def _run(tdse, ctx, s, nsteps):
    p = tdse.Propagator(s.grid(), s.build_potential(), s.dt, precision=s.precision,
                        mask=s.build_mask(), context=ctx)
    p.upload(s.build_psi0())
    p.step(nsteps)
    obs = p.observe()
    psi = p.download()
    p.close()
    return psi, obs


def test_plans_driven_from_many_threads(tdse):
    solves = _solves()
    nsteps = [20 + 7 * i for i in range(len(solves))]
    print("case", [(s.n, s.precision, k) for s, k in zip(solves, nsteps)], flush=True)
    alone_ctx = tdse.Context()
    want = [_run(tdse, alone_ctx, s, k) for s, k in zip(solves, nsteps)]
    alone_ctx.close()

    ctx = tdse.Context()
    got = [None] * len(solves)
    errors = []

    def worker(i):
        try:
            for _ in range(2):  # create, step, observe, destroy, twice
                got[i] = _run(tdse, ctx, solves[i], nsteps[i])
        except Exception as e:  # reported below, on the main thread
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(solves))]
    for th in threads:
        th.start()
    waits = 0
    while any(th.is_alive() for th in threads):
        ctx.sync()
        waits += 1
    for th in threads:
        th.join()
    ctx.sync()
    ctx.close()
    print(f"  {waits} context-wide waits while the workers ran", flush=True)
    assert not errors, errors
    for s, (psi, obs), (psi0, obs0) in zip(solves, got, want):
        assert np.array_equal(psi, psi0), f"{s.n} {s.precision}"
        assert obs == obs0, f"{s.n} {s.precision}"
