"""State uploads and downloads (tdse_state_upload / tdse_state_download /
tdse_state_download_f32).

Transfers go through two pinned buffers of the context in chunks (8 Mi FP32
or 4 Mi FP64 values each); the host threads convert one chunk while the
other is in flight.  A round trip must return the caller's FP64 state bit
for bit (FP64 plans) or rounded once to FP32 (FP32 plans), for sizes of one
partial chunk, whole chunks and several chunks with a remainder, and for
plans of one context transferring from two host threads at once."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def grid3(t, n):
    return t.Grid3D(t.Grid1D(-4.0, 4.0, n[0]), t.Grid1D(-4.0, 4.0, n[1]),
                    t.Grid1D(-4.0, 4.0, n[2]))


def state(n, seed):
    rng = np.random.default_rng(seed)
    shape = (n[2], n[1], n[0])
    # values of every magnitude, so rounding to FP32 is exercised
    mag = np.exp(rng.uniform(-30.0, 5.0, shape))
    return (mag * np.exp(1j * rng.uniform(0.0, 2 * np.pi, shape))).astype(np.complex128)


SHAPES = [
    (16, 8, 4),          # 512 points: part of one chunk
    (256, 128, 128),     # 4 Mi points = 8 Mi values: one FP32 chunk, two FP64 chunks
    (256, 256, 160),     # 10 Mi points: chunks and a remainder in both precisions
]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n", SHAPES, ids=lambda n: "x".join(map(str, n)))
def test_round_trip(tdse, n, prec):
    print(f"case round trip n={n} prec={prec}", flush=True)
    psi = state(n, 7)
    p = tdse.Propagator(grid3(tdse, n), np.zeros(int(np.prod(n))), 0.01, precision=prec)
    p.upload(psi)
    got = p.download()
    want = psi if prec == "fp64" else psi.astype(np.complex64).astype(np.complex128)
    assert np.array_equal(got.view(np.float64), want.view(np.float64))
    # a download into a caller's array, and a second upload over the first
    psi2 = state(n, 8)
    p.upload(psi2)
    out = np.full(psi2.shape, np.nan + 0j)
    p.download(out)
    want2 = psi2 if prec == "fp64" else psi2.astype(np.complex64).astype(np.complex128)
    assert np.array_equal(out.view(np.float64), want2.view(np.float64))
    p.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n", SHAPES, ids=lambda n: "x".join(map(str, n)))
def test_download_f32(tdse, n, prec):
    """tdse_state_download_f32 (the reference engine's float readback): the
    FP32 state as stored, the FP64 state rounded to nearest, i.e. bit for bit
    the FP64 download rounded to complex64, over one or several chunks."""
    print(f"case download f32 n={n} prec={prec}", flush=True)
    p = tdse.Propagator(grid3(tdse, n), np.zeros(int(np.prod(n))), 0.01, precision=prec)
    p.upload(state(n, 9))
    p.step(2)
    got = p.download_f32()
    want = p.download().astype(np.complex64)
    assert got.dtype == np.complex64 and got.shape == want.shape
    assert np.array_equal(got.view(np.float32), want.view(np.float32))
    p.close()


def test_transfers_from_two_threads(tdse):
    """Two plans of one context, each uploaded and read back by its own host
    thread, many times: the shared pinned buffers serialise the transfers."""
    ctx = tdse.Context()
    n = (128, 128, 96)
    plans = [tdse.Propagator(grid3(tdse, n), np.zeros(int(np.prod(n))), 0.01,
                             precision=prec, context=ctx) for prec in ("fp32", "fp64")]
    errors = []

    def worker(k):
        p = plans[k]
        for it in range(6):
            psi = state(n, 100 * k + it)
            p.upload(psi)
            got = p.download()
            want = psi if k == 1 else psi.astype(np.complex64).astype(np.complex128)
            if not np.array_equal(got.view(np.float64), want.view(np.float64)):
                errors.append((k, it))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for p in plans:
        p.close()
    ctx.close()
    assert not errors
