#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ from the reference core.

The reference core is compiled from its own sources by oracle/Makefile
(target "ref", oracle/_ref/libsesref.so); this script calls it through
oracle/ref_capi.cpp and stores inputs and outputs as .npz (no pickles):

  fft.npz           fft/ifft of seeded vectors (1D n=1024, 3D 8x16x4)
  step_small.npz    SplitOperator3D 16x8x32 (mass 0.38, mask) and
                    SplitOperator1D 1024 (mask), a few steps
  ho1d.npz          1D oscillator workload (seed 1): 65536 points, 2000 steps
  hydrogen128.npz   hydrogen-like workload (seed 2): 128^3, 1000 steps
  barrier256.npz    barrier workload (seed 3) in FP64: 256^3, 2000 steps

The large runs keep a sub-lattice of the final state around the packet plus
the observables of the full state, so the files stay small.  Inputs are
rebuilt from the seed by shoeridingerequationsolver_amd.workloads, whose
builders equal the reference builders bit for bit (tests/test_abi.py).

  batch500.npz      the uneven batch workload (seed 4, 500 steps): one solve per
                    grid shape (64^3, 96^3, 128^3, 160^3, 512^2, 1024^2, 65536)
  served.npz        the kept requests of a seeded served stream (1024^2, 192^3,
                    256^3 in both precisions and three smaller ones)

Power-of-two grids run the reference core itself; 96^3, 160^3 and 192^3 (the
reference throws for them, core/src/fft.ixx:42-44) run the oracle's
restatement with its mixed-radix transform, pinned by the brute-force DFT.
The solves of batch500.npz and served.npz are chosen by tests/golden/traces.py.

usage: python tests/golden/make_golden.py [small|ho1d|hydrogen|barrier|batch|served|all]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from shoeridingerequationsolver_amd import workloads  # noqa: E402

R = orc.Ref


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)", flush=True)


def small():
    rng = np.random.default_rng(20260101)
    x1024 = rng.standard_normal(1024) + 1j * rng.standard_normal(1024)
    x3 = rng.standard_normal((4, 16, 8)) + 1j * rng.standard_normal((4, 16, 8))
    save("fft.npz", x1024=x1024, fft1024=R.fft1d(x1024), ifft1024=R.ifft1d(x1024),
         x3=x3, fft3=R.fft3d(x3), ifft3=R.ifft3d(x3))
    dims, lo, hi = (16, 8, 32), (-5.0, -4.0, -8.0), (5.0, 4.0, 8.0)
    v = R.regularized_coulomb_3d(dims, lo, hi, 1.0, (0.0, 0.0, 0.0))
    mask = R.absorbing_mask_3d(dims, lo, hi, 2.0)
    psi0 = R.gaussian_3d(dims, lo, hi, (0.5, 0.0, 1.0), (1.0, 1.2, 2.0), (0.7, 0.0, -0.3))
    psi0 = psi0.reshape(dims[::-1])
    psi = R.so3d_evolve(v, dims, lo, hi, 0.01, psi0, 7, mass=0.38, mask=mask)
    lo1, hi1, n1 = -40.0, 40.0, 1024
    v1 = R.harmonic_1d(lo1, hi1, n1, 0.1, 0.0)
    m1 = R.absorbing_mask_1d(lo1, hi1, n1, 8.0)
    p10 = R.gaussian_1d(lo1, hi1, n1, -15.0, 1.0, 2.0)
    p1 = R.so1d_evolve(v1, lo1, hi1, 0.05, p10, 20, mask=m1)
    save("step_small.npz", dims=np.array(dims), lo=np.array(lo), hi=np.array(hi), v=v, mask=mask,
         psi0=psi0, psi=psi, dt=0.01, steps=7, mass=0.38, v1=v1, mask1=m1, psi1_0=p10, psi1=p1,
         lo1=lo1, hi1=hi1, dt1=0.05, steps1=20)


def _observe3(s, psi, v):
    dims, lo, hi = s.n, s.lo, s.hi
    cell = np.prod([(h - l) / n for l, h, n in zip(lo, hi, dims)])
    return dict(norm=orc.norm_sq(psi, cell),
                position=orc.mean_position_3d(psi, dims, lo, hi),
                energy=R.mean_energy_3d(psi, dims, lo, hi, v),
                max_abs=np.abs(psi).max(),
                transmitted=orc.probability_in_slab_3d(psi, dims, lo, hi, 0, 2.0, 1e300))


def ho1d():
    s = workloads.ho1d(seed=1)
    v, psi0 = s.build_potential(), s.build_psi0()
    t0 = time.time()
    psi = R.so1d_evolve(v, s.lo[0], s.hi[0], s.dt, psi0, s.steps)
    print(f"ho1d reference run {time.time() - t0:.1f} s", flush=True)
    idx = np.arange(0, s.n[0], 16)
    save("ho1d.npz", seed=s.seed, steps=s.steps, idx=idx, psi=psi[idx],
         norm=orc.norm_sq(psi, (s.hi[0] - s.lo[0]) / s.n[0]),
         position=R.mean_position_1d(psi, s.lo[0], s.hi[0]),
         energy=R.mean_energy_1d(psi, s.lo[0], s.hi[0], v), max_abs=np.abs(psi).max())


def _sublattice(shape, step, center, half):
    """Flat indices of every `step`-th point within `half` points of `center`."""
    axes = [np.arange(max(0, c - half), min(n, c + half), step) for n, c in zip(shape, center)]
    k, j, i = np.meshgrid(axes[0], axes[1], axes[2], indexing="ij")
    nz, ny, nx = shape
    return (i + nx * (j + ny * k)).ravel()


def hydrogen():
    s = workloads.hydrogen(seed=2)
    v, psi0 = s.build_potential(), s.build_psi0()
    t0 = time.time()
    psi = R.so3d_evolve(v, s.n, s.lo, s.hi, s.dt, psi0, s.steps)
    print(f"hydrogen reference run {time.time() - t0:.1f} s", flush=True)
    idx = _sublattice(psi.shape, 3, (64, 64, 64), 30)
    save("hydrogen128.npz", seed=s.seed, steps=s.steps, idx=idx, psi=psi.ravel()[idx],
         **_observe3(s, psi, v))


def barrier():
    s = workloads.barrier(seed=3, precision="fp64")
    v, mask, psi0 = s.build_potential(), s.build_mask(), s.build_psi0()
    t0 = time.time()
    psi = R.so3d_evolve(v, s.n, s.lo, s.hi, s.dt, psi0, s.steps, mask=mask)
    print(f"barrier reference run {time.time() - t0:.1f} s", flush=True)
    # the packet ends near x ~ +10 (transmitted) and x ~ -35 (reflected)
    idx = _sublattice(psi.shape, 4, (128, 128, 128), 40)
    line = (np.arange(256) + 256 * (128 + 256 * 128))
    save("barrier256.npz", seed=s.seed, steps=s.steps, idx=idx, psi=psi.ravel()[idx],
         line_idx=line, line=psi.ravel()[line], **_observe3(s, psi, v))


def _evolve(s):
    """FP64 reference run of one solve: (psi in the field's shape, v)."""
    v, mask, psi0 = s.build_potential(), s.build_mask(), s.build_psi0()
    dims, lo, hi = s.n, s.lo, s.hi
    if dims[1] == dims[2] == 1:
        return R.so1d_evolve(v, lo[0], hi[0], s.dt, psi0, s.steps, mask=mask), v
    pow2 = all(d & (d - 1) == 0 for d in dims)
    psi0 = psi0.reshape(dims[::-1])
    if pow2:
        return R.so3d_evolve(v, dims, lo, hi, s.dt, psi0, s.steps, mask=mask), v
    return orc.so3d_evolve(v, dims, lo, hi, s.dt, psi0, s.steps, mask=mask, any_size=True), v


def _record(out, key, s, psi, v):
    from traces import sample_indices
    dims, lo, hi = s.n, s.lo, s.hi
    flat = psi.ravel()
    idx = sample_indices(dims)
    cell = np.prod([(h - l) / n for l, h, n in zip(lo, hi, dims)])
    out[f"{key}_idx"] = idx
    out[f"{key}_psi"] = flat[idx]
    out[f"{key}_norm"] = orc.norm_sq(flat, cell)
    out[f"{key}_max_abs"] = np.abs(flat).max()
    if dims[1] == dims[2] == 1:
        out[f"{key}_position"] = np.array([R.mean_position_1d(flat, lo[0], hi[0]), 0.0, 0.0])
        out[f"{key}_energy"] = R.mean_energy_1d(flat, lo[0], hi[0], v)
    else:
        f3 = flat.reshape(dims[::-1])
        out[f"{key}_position"] = np.asarray(orc.mean_position_3d(f3, dims, lo, hi))
        out[f"{key}_energy"] = orc.mean_energy_3d(f3, dims, lo, hi, v)


def batch():
    from traces import BATCH_COUNT, BATCH_SEED, BATCH_STEPS, batch_golden, tag
    out = {"seed": BATCH_SEED, "count": BATCH_COUNT, "steps": BATCH_STEPS}
    for i, s in batch_golden():
        t0 = time.time()
        psi, v = _evolve(s)
        print(f"batch {s.name} {s.n} {s.potential} reference run {time.time() - t0:.1f} s",
              flush=True)
        _record(out, tag(s.n), s, psi, v)
        out[f"{tag(s.n)}_index"] = i
    save("batch500.npz", **out)


def served():
    from traces import SERVED_SEED, served_golden
    out = {"seed": SERVED_SEED}
    for t, s in served_golden():
        t0 = time.time()
        psi, v = _evolve(s)
        print(f"served {s.name} {s.n} {s.precision} {s.potential} mask={s.mask_width} "
              f"steps={s.steps} reference run {time.time() - t0:.1f} s", flush=True)
        _record(out, s.name, s, psi, v)
    save("served.npz", **out)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "small"
    if not orc.ref_available():
        sys.exit("oracle/_ref/libsesref.so is missing: run `make -C oracle ref`")
    sys.path.insert(0, HERE)
    for name, fn in (("small", small), ("ho1d", ho1d), ("hydrogen", hydrogen),
                     ("barrier", barrier), ("batch", batch), ("served", served)):
        if what in (name, "all"):
            fn()
