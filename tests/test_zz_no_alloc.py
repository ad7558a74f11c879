"""tdse_observe_noalloc (the entry point that allocates no device memory)
against tdse_observe on the same states, through the C ABI: two calls in a
row, the second on a smaller state, both precisions.  The reduction runs the
same blocks and summation order, so the sums are bitwise equal and the
ratios equal to the last bit of FP64 division and square root."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipStreamSynchronize.argtypes = [C.c_void_p]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), n) == 0
        return p

    def to_host(self, dst, src, n):
        assert self.lib.hipMemcpy(C.cast(dst, C.c_void_p), src, n, 2) == 0  # device to host


def _observe_noalloc(t, hip, prop, out_dev):
    from shoeridingerequationsolver_amd import capi
    L = capi.lib()
    n, lo, hi = t._axes(prop.grid)
    capi.check(L.tdse_observe_noalloc(C.c_void_p(prop.device_ptr()), None, prop.precision,
                                      capi.ivec(n), capi.dvec(lo), capi.dvec(hi), out_dev,
                                      C.c_void_p(prop.stream())))
    assert hip.lib.hipStreamSynchronize(C.c_void_p(prop.stream())) == 0
    o = capi.Observables()
    hip.to_host(C.byref(o), out_dev, C.sizeof(o))
    return o


CASES = [((64, 32, 96), (32, 16, 8)), ((256, 128, 1), (96, 64, 1)), ((65536, 1, 1), (1024, 1, 1))]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("shapes", CASES, ids=["3d", "2d", "1d"])
def test_noalloc_equals_observe(tdse, prec, shapes):
    from shoeridingerequationsolver_amd import capi
    hip = _Hip()
    out_dev = hip.malloc(C.sizeof(capi.Observables))
    try:
        for rep, shape in enumerate(shapes):  # the second call on a smaller state
            dims = sum(1 for v in shape if v > 1)
            half = {3: 8.0, 2: 12.0, 1: 50.0}[dims]
            axes = [tdse.Grid1D(-half, half, v) if v > 1 else tdse.Grid1D(0.0, 1.0, 1)
                    for v in shape]
            grid = axes[0] if dims == 1 else tdse.Grid3D(*axes)
            rng = np.random.default_rng(100 + rep)
            psi = (rng.standard_normal(int(np.prod(shape))) +
                   1j * rng.standard_normal(int(np.prod(shape)))).reshape(
                       (shape[0],) if dims == 1 else shape[::-1])
            psi *= np.exp(-0.5 * np.linspace(-3, 3, psi.size) ** 2).reshape(psi.shape)
            p = tdse.Propagator(grid, np.zeros(psi.size), 0.01, precision=prec)
            p.upload(psi)
            want = p.observe(capi.TDSE_OBS_NORM | capi.TDSE_OBS_POSITION)
            got = _observe_noalloc(tdse, hip, p, out_dev)
            print(f"case {prec} {shape} call {rep}: norm {got.norm:.17g} vs {want['norm']:.17g}",
                  flush=True)
            assert got.sum_abs2 == want["sum_abs2"]
            assert got.norm == want["norm"]
            assert tuple(got.position) == want["position"]
            np.testing.assert_allclose(tuple(got.sigma), want["sigma"], rtol=4e-16, atol=0)
            assert got.potential == want["potential"] == 0.0
            assert got.kinetic == got.energy == 0.0
            p.close()
    finally:
        hip.lib.hipFree(out_dev)


def test_noalloc_potential_term(tdse):
    """<V> through a caller-held potential array equals tdse_observe's."""
    from shoeridingerequationsolver_amd import capi
    hip = _Hip()
    shape = (64, 32, 16)
    grid = tdse.Grid3D(tdse.Grid1D(-8, 8, 64), tdse.Grid1D(-6, 6, 32), tdse.Grid1D(-4, 4, 16))
    v = tdse.harmonic_potential(grid, 0.7, (0.5, 0.0, -0.25))
    psi = tdse.gaussian_wavepacket(grid, (1.0, 0.5, 0.0), (1.2, 1.0, 0.8), (0.3, 0.0, 0.0))
    L = capi.lib()
    for prec, dt in (("fp64", np.float64), ("fp32", np.float32)):
        p = tdse.Propagator(grid, v, 0.01, precision=prec)
        p.upload(psi)
        want = p.observe(capi.TDSE_OBS_NORM | capi.TDSE_OBS_POSITION | capi.TDSE_OBS_ENERGY)
        vh = np.ascontiguousarray(v.astype(dt))
        v_dev = hip.malloc(vh.nbytes)
        out_dev = hip.malloc(C.sizeof(capi.Observables))
        assert hip.lib.hipMemcpy(v_dev, vh.ctypes.data_as(C.c_void_p), vh.nbytes, 1) == 0
        n, lo, hi = tdse._axes(grid)
        capi.check(L.tdse_observe_noalloc(C.c_void_p(p.device_ptr()), v_dev, p.precision,
                                          capi.ivec(n), capi.dvec(lo), capi.dvec(hi), out_dev,
                                          C.c_void_p(p.stream())))
        assert hip.lib.hipStreamSynchronize(C.c_void_p(p.stream())) == 0
        o = capi.Observables()
        hip.to_host(C.byref(o), out_dev, C.sizeof(o))
        print(f"case potential {prec}: <V> {o.potential:.17g} vs {want['potential']:.17g}",
              flush=True)
        assert o.potential == want["potential"]
        assert o.norm == want["norm"]
        hip.lib.hipFree(v_dev)
        hip.lib.hipFree(out_dev)
        p.close()


def test_noalloc_rejects_bad_arguments(tdse):
    from shoeridingerequationsolver_amd import capi
    L = capi.lib()
    o = capi.Observables()
    n = capi.ivec((4, 4, 4))
    lo, hi = capi.dvec((0, 0, 0)), capi.dvec((1, 1, 1))
    assert L.tdse_observe_noalloc(None, None, 0, n, lo, hi, C.byref(o), None) == capi.TDSE_E_INVALID
    big = capi.ivec((1024, 1024, 1024))
    # a non-null stand-in pointer: the size check comes before any launch
    assert L.tdse_observe_noalloc(C.c_void_p(256), None, 0, big, lo, hi, C.byref(o),
                                  None) == capi.TDSE_E_UNSUPPORTED
    assert L.tdse_observe_noalloc(C.c_void_p(256), None, 7, n, lo, hi, C.byref(o),
                                  None) == capi.TDSE_E_INVALID
