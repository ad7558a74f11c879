"""The seeded solves whose reference results are stored as golden vectors
for the two multi-solve workloads (shared by make_golden.py and the GPU
tests, so both rebuild exactly the same inputs):

* batch_golden(): the uneven batch of 48 solves (seed 4, 500 steps each),
  one instance per grid shape of workloads.BATCH_SHAPES (64^3, 96^3, 128^3,
  160^3, 512^2, 1024^2, 65536), the first of that shape in the seeded list;
* served_golden(): a Poisson request stream (seed 41, 400 requests/s, steps
  U(100, 160)) from which the first request of each (shape, precision) in
  SERVED_KEYS is kept, in arrival order: 1024^2, 192^3 and 256^3 in both
  precisions, plus 65536 FP32, 64^3 FP64 and 128^3 FP32.

sample_indices() picks the stored points: a regular sub-lattice of about
2000 points over the whole grid plus (2D, 3D) the full x line through the
centre."""
import math

import numpy as np

BATCH_SEED, BATCH_COUNT, BATCH_STEPS = 4, 48, 500
SERVED_SEED, SERVED_RATE, SERVED_MAX_STEPS = 41, 400.0, 160
SERVED_KEYS = [((1024, 1024, 1), "fp32"), ((1024, 1024, 1), "fp64"),
               ((192, 192, 192), "fp32"), ((192, 192, 192), "fp64"),
               ((256, 256, 256), "fp32"), ((256, 256, 256), "fp64"),
               ((65536, 1, 1), "fp32"), ((64, 64, 64), "fp64"), ((128, 128, 128), "fp32")]


def tag(n):
    return "x".join(str(v) for v in n)


def batch_golden():
    """[(index in the seeded list, Solve)] one per shape, list order."""
    from shoeridingerequationsolver_amd import workloads
    solves = workloads.batch_instances(seed=BATCH_SEED, count=BATCH_COUNT, steps=BATCH_STEPS)
    seen, out = set(), []
    for i, s in enumerate(solves):
        if s.n not in seen:
            seen.add(s.n)
            out.append((i, s))
    return out


def served_golden():
    """[(arrival_s, Solve)] of the kept requests, in arrival order."""
    from shoeridingerequationsolver_amd import serving
    full = serving.request_stream(seed=SERVED_SEED, count=400, rate=SERVED_RATE,
                                  max_steps=SERVED_MAX_STEPS)
    kept = []
    for key in SERVED_KEYS:
        for t, s in full:
            if (s.n, s.precision) == key:
                kept.append((t, s))
                break
        else:
            raise RuntimeError(f"no request {key} in the seeded stream")
    return sorted(kept, key=lambda x: x[0])


def sample_indices(n, target=2000):
    """Flat (x-fastest) indices of a regular sub-lattice over the grid and of
    the x line through the centre."""
    dims = [v for v in n if v > 1]
    total = int(np.prod(n))
    stride = max(1, int(math.ceil((total / target) ** (1.0 / len(dims)))))
    axes = [np.arange(stride // 2, v, stride) if v > 1 else np.array([0]) for v in n]
    k, j, i = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    lattice = (i + n[0] * (j + n[1] * k)).ravel()
    if len(dims) == 1:
        return lattice
    line = np.arange(n[0]) + n[0] * ((n[1] // 2) + n[1] * (n[2] // 2))
    return np.unique(np.concatenate([lattice, line]))
