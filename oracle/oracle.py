"""ctypes bindings for the parity checker.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module.  It loads

* ``oracle/_build/liborc.so`` -- the C restatement of the reference core
  (``tdse_oracle.c``), and
* ``oracle/_ref/libsesref.so`` -- the reference core itself, compiled from its
  own sources by ``oracle/Makefile`` (present only where it was built).

Complex arrays are numpy ``complex128`` (interleaved re/im, the layout of
``std::complex<double>`` in ``ses::Field1D/Field3D``, core/src/field.ixx:11-66).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_PATH = os.path.join(HERE, "_build", "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libsesref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(ref: bool = False) -> None:
    """Compile the C restatement (and, with ``ref``, the reference core)."""
    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE, "-j8", *targets], check=True)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _ivec(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def _dvec(v):
    return (C.c_double * 3)(*[float(x) for x in v])


_ORC = None
_REF = None


def orc() -> C.CDLL:
    global _ORC
    if _ORC is None:
        if not os.path.exists(ORC_PATH):
            build()
        lib = C.CDLL(ORC_PATH)
        for name in ("orc_norm_sq", "orc_mean_position_1d", "orc_mean_energy_1d",
                     "orc_energy_variance_1d",
                     "orc_probability_in_range_1d", "orc_mean_energy_3d",
                     "orc_probability_in_slab_3d", "orc_probability_in_box_3d",
                     "orc_spacing", "orc_sigma_x_1d",
                     "orc_mean_momentum_1d", "orc_lat2d_energy"):
            getattr(lib, name).restype = C.c_double
        _ORC = lib
    return _ORC


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> C.CDLL:
    global _REF
    if _REF is None:
        lib = C.CDLL(REF_PATH)
        for name in ("ref_norm_sq_1d", "ref_norm_sq_3d", "ref_mean_position_1d",
                     "ref_mean_energy_1d", "ref_probability_in_range_1d", "ref_mean_energy_3d",
                     "ref_energy_variance_1d",
                     "ref_sigma_x_1d", "ref_mean_momentum_1d"):
            getattr(lib, name).restype = C.c_double
        _REF = lib
    return _REF


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


# ---- C restatement ---------------------------------------------------------

def wavenumbers(n, xmin, xmax):
    k = np.empty(n)
    orc().orc_wavenumbers(C.c_int(n), C.c_double(xmin), C.c_double(xmax), _p(k))
    return k


def fft1d(a):
    a = _c128(a).copy()
    if orc().orc_fft1d(_p(a), C.c_size_t(a.size)) != 0:
        raise ValueError("fft: size must be a power of two")
    return a


def ifft1d(a):
    a = _c128(a).copy()
    if orc().orc_ifft1d(_p(a), C.c_size_t(a.size)) != 0:
        raise ValueError("ifft: size must be a power of two")
    return a


def fft3d(a):
    """Forward 3D FFT of an array shaped (nz, ny, nx) (x fastest)."""
    a = _c128(a).copy()
    nz, ny, nx = a.shape
    if orc().orc_fft3d(_p(a), nx, ny, nz) != 0:
        raise ValueError("fft: size must be a power of two")
    return a


def ifft3d(a):
    a = _c128(a).copy()
    nz, ny, nx = a.shape
    if orc().orc_ifft3d(_p(a), nx, ny, nz) != 0:
        raise ValueError("fft: size must be a power of two")
    return a


def fft_any(a, sign=-1):
    a = _c128(a).copy()
    orc().orc_fft_any(_p(a), C.c_size_t(a.size), C.c_int(sign))
    return a


def dft_brute(a, sign=-1):
    a = _c128(a)
    out = np.empty_like(a)
    orc().orc_dft_brute(_p(a), _p(out), C.c_size_t(a.size), C.c_int(sign))
    return out


def fft3d_any(a, sign=-1):
    a = _c128(a).copy()
    nz, ny, nx = a.shape
    orc().orc_fft3d_any(_p(a), nx, ny, nz, C.c_int(sign))
    return a


def so1d_tables(v, xmin, xmax, dt):
    v = np.ascontiguousarray(v, dtype=np.float64)
    n = v.size
    hv = np.empty(n, np.complex128)
    kin = np.empty(n, np.complex128)
    orc().orc_so1d_tables(_p(v), n, C.c_double(xmin), C.c_double(xmax), C.c_double(dt),
                          _p(hv), _p(kin))
    return hv, kin


def so1d_evolve(v, xmin, xmax, dt, psi, nsteps, mask=None):
    """SplitOperator1D::step over nsteps; with ``mask``, psi *= mask after each step."""
    hv, kin = so1d_tables(v, xmin, xmax, dt)
    psi = _c128(psi).copy()
    n = psi.size
    if mask is None:
        if orc().orc_so1d_step(_p(psi), n, _p(hv), _p(kin), nsteps) != 0:
            raise ValueError("size must be a power of two")
    else:
        mask = np.ascontiguousarray(mask, dtype=np.float64)
        for _ in range(nsteps):
            if orc().orc_so1d_step(_p(psi), n, _p(hv), _p(kin), 1) != 0:
                raise ValueError("size must be a power of two")
            orc().orc_apply_mask(_p(psi), _p(mask), C.c_size_t(n))
    return psi


def so3d_tables(v, dims, lo, hi, dt, mass=1.0):
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    hv = np.empty(v.size, np.complex128)
    kin = np.empty(v.size, np.complex128)
    orc().orc_so3d_tables(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(dt),
                          C.c_double(mass), _p(hv), _p(kin))
    return hv, kin


def so3d_evolve(v, dims, lo, hi, dt, psi, nsteps, mass=1.0, mask=None, any_size=False):
    """SplitOperator3D::step; ``any_size`` switches to the mixed-radix FFT."""
    hv, kin = so3d_tables(v, dims, lo, hi, dt, mass)
    psi = _c128(psi).copy()
    shape = psi.shape
    psi = psi.ravel()
    d = _ivec(dims)
    step = orc().orc_so3d_step_any if any_size else orc().orc_so3d_step
    if mask is None:
        if step(_p(psi), d, _p(hv), _p(kin), nsteps) != 0:
            raise ValueError("size must be a power of two")
    else:
        mask = np.ascontiguousarray(mask, dtype=np.float64).ravel()
        for _ in range(nsteps):
            if step(_p(psi), d, _p(hv), _p(kin), 1) != 0:
                raise ValueError("size must be a power of two")
            orc().orc_apply_mask(_p(psi), _p(mask), C.c_size_t(psi.size))
    return psi.reshape(shape)


def sigma_x_1d(psi, xmin, xmax):
    psi = _c128(psi)
    return orc().orc_sigma_x_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax))


def mean_momentum_1d(psi, xmin, xmax):
    psi = _c128(psi)
    return orc().orc_mean_momentum_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax))


def sigma_position_3d(psi, dims, lo, hi):
    psi = _c128(psi).ravel()
    out = np.empty(3)
    orc().orc_sigma_position_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(out))
    return tuple(out)


def mean_momentum_3d(psi, dims, lo, hi):
    psi = _c128(psi).ravel()
    out = np.empty(3)
    orc().orc_mean_momentum_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(out))
    return tuple(out)


def it1d_tables(v, xmin, xmax, dtau):
    """ImaginaryTimePropagator1D weights (core/src/imaginary_time.ixx:24-40)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    hw = np.empty(v.size)
    kw = np.empty(v.size)
    orc().orc_it1d_tables(_p(v), v.size, C.c_double(xmin), C.c_double(xmax), C.c_double(dtau),
                          _p(hw), _p(kw))
    return hw, kw


def it1d_relax(v, xmin, xmax, dtau, psi, nsteps):
    """ImaginaryTimePropagator1D::relax; returns (psi, norm of the last step
    before its renormalisation)."""
    hw, kw = it1d_tables(v, xmin, xmax, dtau)
    psi = _c128(psi).copy()
    last = C.c_double(0.0)
    dx = (xmax - xmin) / psi.size
    if orc().orc_it1d_relax(_p(psi), psi.size, C.c_double(dx), _p(hw), _p(kw), nsteps,
                            C.byref(last)) != 0:
        raise ValueError("size must be a power of two")
    return psi, last.value


def it3d_tables(v, dims, lo, hi, dtau, mass=1.0):
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    hw = np.empty(v.size)
    kw = np.empty(v.size)
    orc().orc_it3d_tables(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(dtau),
                          C.c_double(mass), _p(hw), _p(kw))
    return hw, kw


def it3d_relax(v, dims, lo, hi, dtau, psi, nsteps, mass=1.0, lower=(), any_size=False):
    """ImaginaryTimePropagator3D::relax, or relax_deflated against `lower`
    (core/src/imaginary_time.ixx:100-134); returns (psi, last pre-renorm norm)."""
    hw, kw = it3d_tables(v, dims, lo, hi, dtau, mass)
    psi = _c128(psi).copy()
    shape = psi.shape
    psi = psi.ravel()
    low = [_c128(q).ravel() for q in lower]
    arr = (C.c_void_p * max(1, len(low)))(*[q.ctypes.data for q in low])
    dv = 1.0
    for a in range(3):
        dv *= (hi[a] - lo[a]) / dims[a]
    last = C.c_double(0.0)
    if orc().orc_it3d_relax(_p(psi), _ivec(dims), C.c_double(dv), _p(hw), _p(kw), arr, len(low),
                            nsteps, 1 if any_size else 0, C.byref(last)) != 0:
        raise ValueError("size must be a power of two")
    return psi.reshape(shape), last.value


def inner_product(a, b, dv):
    a = _c128(a).ravel()
    b = _c128(b).ravel()
    out = np.empty(2)
    orc().orc_inner_product(_p(a), _p(b), C.c_size_t(a.size), C.c_double(dv), _p(out))
    return complex(out[0], out[1])


def soft_coulomb_1d(xmin, xmax, n, z, a, x0=0.0):
    v = np.empty(n)
    orc().orc_soft_coulomb_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(z), C.c_double(a),
                              C.c_double(x0), _p(v))
    return v


def soft_coulomb_3d(dims, lo, hi, z, a, c):
    v = np.empty(int(np.prod(dims)))
    orc().orc_soft_coulomb_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(z), C.c_double(a),
                              _dvec(c), _p(v))
    return v


def driven_evolve(v, dims, lo, hi, dt, psi, nsteps, axis, amplitude, omega, t0=0.0, mass=1.0,
                  any_size=False):
    """ses::driven_step (core/src/drive.ixx:52-66) with SplitOperator3D tables."""
    hv, kin = so3d_tables(v, dims, lo, hi, dt, mass)
    psi = _c128(psi).copy()
    shape = psi.shape
    psi = psi.ravel()
    if orc().orc_driven_step(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(hv), _p(kin),
                             _dvec(axis), C.c_double(amplitude), C.c_double(omega),
                             C.c_double(t0), C.c_double(dt), nsteps,
                             1 if any_size else 0) != 0:
        raise ValueError("size must be a power of two")
    return psi.reshape(shape)


def rotate_axis(psi, dims, lo, hi, axis, theta):
    psi = _c128(psi).copy()
    if orc().orc_rotate_axis(_p(psi.ravel()), _ivec(dims), _dvec(lo), _dvec(hi), axis,
                             C.c_double(theta)) != 0:
        raise ValueError("size must be a power of two")
    return psi


def magnetic_veff(v, dims, lo, hi, bfield, axis):
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    out = np.empty_like(v)
    orc().orc_magnetic_veff(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(bfield), axis,
                            _p(out))
    return out


def magnetic_evolve(v, dims, lo, hi, dt, bfield, axis, psi, nsteps, mask=None):
    """MagneticPropagator3D::step (core/src/magnetic.ixx:34-46); with ``mask``,
    psi *= mask after each step."""
    veff = magnetic_veff(v, dims, lo, hi, bfield, axis)
    hv, kin = so3d_tables(veff, dims, lo, hi, dt, 1.0)
    psi = _c128(psi).copy()
    shape = psi.shape
    psi = psi.ravel()
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64).ravel()
    for _ in range(nsteps if m is not None else 1):
        k = 1 if m is not None else nsteps
        if orc().orc_magnetic_step(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(hv), _p(kin),
                                   C.c_double(dt), C.c_double(bfield), axis, k) != 0:
            raise ValueError("size must be a power of two")
        if m is not None:
            orc().orc_apply_mask(_p(psi), _p(m), C.c_size_t(psi.size))
    return psi.reshape(shape)


def lat2d_links(dims, lo, hi, links=0, a=0.0, xs=0.0, ys=0.0, cut_up=True):
    """x-link phases of PeierlsLattice2D (1 everywhere, a uniform field or a
    solenoid), interleaved complex, row-major (j, i)."""
    nx, ny = dims[0], dims[1]
    lx = np.empty(nx * ny, np.complex128)
    if links == 1:
        orc().orc_lat2d_uniform_field(_p(lx), nx, ny, C.c_double(lo[0]), C.c_double(hi[0]),
                                      C.c_double(lo[1]), C.c_double(hi[1]), C.c_double(a))
    else:
        orc().orc_lat2d_solenoid(_p(lx), nx, ny, C.c_double(lo[0]), C.c_double(hi[0]),
                                 C.c_double(lo[1]), C.c_double(hi[1]),
                                 C.c_double(a if links == 2 else 0.0),
                                 C.c_double(xs if links == 2 else 1e300),
                                 C.c_double(ys), 1 if cut_up else 0)
    return lx


def lat2d_evolve(v, dims, lo, hi, dt, psi, nsteps, relax=False, links=0, a=0.0, xs=0.0, ys=0.0,
                 cut_up=True):
    """PeierlsLattice2D::step / relax; returns (psi, energy of the result)."""
    nx, ny = dims[0], dims[1]
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    hx = (hi[0] - lo[0]) / nx
    hy = (hi[1] - lo[1]) / ny
    consts = np.empty(14)
    hv = np.empty(nx * ny, np.complex128)
    rhv = np.empty(nx * ny)
    orc().orc_lat2d_tables(_p(v), nx, ny, C.c_double(hx), C.c_double(hy), C.c_double(dt),
                           _p(consts), _p(hv), _p(rhv))
    lx = lat2d_links(dims, lo, hi, links, a, xs, ys, cut_up)
    psi = _c128(psi).copy()
    shape = psi.shape
    psi = psi.ravel()
    dv = hx * hy * (hi[2] - lo[2]) / dims[2]
    orc().orc_lat2d_evolve(_p(psi), nx, ny, _p(consts), _p(hv), _p(rhv), _p(lx),
                           1 if relax else 0, C.c_double(dv), nsteps)
    e = orc().orc_lat2d_energy(_p(psi), nx, ny, _p(v), _p(consts), _p(lx))
    return psi.reshape(shape), e


def norm_sq(psi, cell_volume):
    psi = _c128(psi)
    return orc().orc_norm_sq(_p(psi), C.c_size_t(psi.size), C.c_double(cell_volume))


def harmonic_1d(xmin, xmax, n, omega, x0=0.0):
    v = np.empty(n)
    orc().orc_harmonic_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(omega),
                          C.c_double(x0), _p(v))
    return v


def harmonic_3d(dims, lo, hi, omega, c):
    v = np.empty(int(np.prod(dims)))
    orc().orc_harmonic_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(omega), _dvec(c), _p(v))
    return v


def barrier_1d(xmin, xmax, n, v0, x_lo, x_hi):
    v = np.empty(n)
    orc().orc_barrier_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(v0),
                         C.c_double(x_lo), C.c_double(x_hi), _p(v))
    return v


def barrier_3d(dims, lo, hi, v0, x_lo, x_hi):
    v = np.empty(int(np.prod(dims)))
    orc().orc_barrier_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(v0), C.c_double(x_lo),
                         C.c_double(x_hi), _p(v))
    return v


def regularized_coulomb_3d(dims, lo, hi, z, c):
    v = np.empty(int(np.prod(dims)))
    orc().orc_regularized_coulomb_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(z), _dvec(c),
                                     _p(v))
    return v


def coulomb_floor_3d(dims, lo, hi, z, c):
    v = np.empty(int(np.prod(dims)))
    orc().orc_coulomb_floor_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(z), _dvec(c), _p(v))
    return v


def absorbing_mask_1d(xmin, xmax, n, width):
    m = np.empty(n)
    orc().orc_absorbing_mask_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(width), _p(m))
    return m


def absorbing_mask_3d(dims, lo, hi, width):
    m = np.empty(int(np.prod(dims)))
    orc().orc_absorbing_mask_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(width), _p(m))
    return m


def gaussian_1d(xmin, xmax, n, x0, sigma, k0):
    psi = np.empty(n, np.complex128)
    orc().orc_gaussian_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(x0),
                          C.c_double(sigma), C.c_double(k0), _p(psi))
    return psi


def gaussian_3d(dims, lo, hi, r0, sigma, k0):
    psi = np.empty(int(np.prod(dims)), np.complex128)
    orc().orc_gaussian_3d(_ivec(dims), _dvec(lo), _dvec(hi), _dvec(r0), _dvec(sigma), _dvec(k0),
                          _p(psi))
    return psi


def mean_position_1d(psi, xmin, xmax):
    psi = _c128(psi)
    return orc().orc_mean_position_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax))


def mean_energy_1d(psi, xmin, xmax, v):
    psi = _c128(psi)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return orc().orc_mean_energy_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax), _p(v))


def energy_variance_1d(psi, xmin, xmax, v):
    """energy_variance (core/src/observables.ixx:92-119), residual form."""
    psi = _c128(psi)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return orc().orc_energy_variance_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax),
                                        _p(v))


def probability_in_range_1d(psi, xmin, xmax, a, b):
    psi = _c128(psi)
    return orc().orc_probability_in_range_1d(_p(psi), psi.size, C.c_double(xmin),
                                             C.c_double(xmax), C.c_double(a), C.c_double(b))


def mean_position_3d(psi, dims, lo, hi):
    psi = _c128(psi)
    out = np.empty(3)
    orc().orc_mean_position_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(out))
    return out


def mean_energy_3d(psi, dims, lo, hi, v, mass=1.0):
    psi = _c128(psi)
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    return orc().orc_mean_energy_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(v),
                                    C.c_double(mass))


def probability_in_slab_3d(psi, dims, lo, hi, axis, a, b):
    psi = _c128(psi)
    return orc().orc_probability_in_slab_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi),
                                            C.c_int(axis), C.c_double(a), C.c_double(b))


def probability_in_box_3d(psi, dims, lo, hi, a, b):
    psi = _c128(psi)
    return orc().orc_probability_in_box_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi),
                                           _dvec(a), _dvec(b))


# ---- the reference core (oracle/_ref) --------------------------------------

class Ref:
    """Thin wrappers over ``libsesref.so`` with the same signatures as above."""

    @staticmethod
    def fft1d(a):
        a = _c128(a).copy()
        if ref().ref_fft1d(_p(a), C.c_size_t(a.size)) != 0:
            raise ValueError("fft: size must be a power of two")
        return a

    @staticmethod
    def ifft1d(a):
        a = _c128(a).copy()
        if ref().ref_ifft1d(_p(a), C.c_size_t(a.size)) != 0:
            raise ValueError("ifft: size must be a power of two")
        return a

    @staticmethod
    def fft3d(a):
        a = _c128(a).copy()
        nz, ny, nx = a.shape
        if ref().ref_fft3d(_p(a), _ivec((nx, ny, nz))) != 0:
            raise ValueError("fft: size must be a power of two")
        return a

    @staticmethod
    def ifft3d(a):
        a = _c128(a).copy()
        nz, ny, nx = a.shape
        if ref().ref_ifft3d(_p(a), _ivec((nx, ny, nz))) != 0:
            raise ValueError("fft: size must be a power of two")
        return a

    @staticmethod
    def wavenumbers(n, xmin, xmax):
        k = np.empty(n)
        ref().ref_wavenumbers(n, C.c_double(xmin), C.c_double(xmax), _p(k))
        return k

    @staticmethod
    def so1d_evolve(v, xmin, xmax, dt, psi, nsteps, mask=None):
        v = np.ascontiguousarray(v, dtype=np.float64)
        psi = _c128(psi).copy()
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
        rc = ref().ref_so1d_evolve(_p(v), v.size, C.c_double(xmin), C.c_double(xmax),
                                   C.c_double(dt), _p(psi), nsteps,
                                   None if m is None else _p(m))
        if rc != 0:
            raise ValueError("size must be a power of two")
        return psi

    @staticmethod
    def so3d_tables(v, dims, lo, hi, dt, mass=1.0):
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        hv = np.empty(v.size, np.complex128)
        kin = np.empty(v.size, np.complex128)
        ref().ref_so3d_tables(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(dt),
                              C.c_double(mass), _p(hv), _p(kin))
        return hv, kin

    @staticmethod
    def so3d_evolve(v, dims, lo, hi, dt, psi, nsteps, mass=1.0, mask=None):
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        psi = _c128(psi).copy()
        shape = psi.shape
        psi = psi.ravel()
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64).ravel()
        rc = ref().ref_so3d_evolve(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(dt),
                                   C.c_double(mass), _p(psi), nsteps,
                                   None if m is None else _p(m))
        if rc != 0:
            raise ValueError("size must be a power of two")
        return psi.reshape(shape)

    @staticmethod
    def harmonic_1d(xmin, xmax, n, omega, x0=0.0):
        v = np.empty(n)
        ref().ref_harmonic_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(omega),
                              C.c_double(x0), _p(v))
        return v

    @staticmethod
    def harmonic_3d(dims, lo, hi, omega, c):
        v = np.empty(int(np.prod(dims)))
        ref().ref_harmonic_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(omega), _dvec(c),
                              _p(v))
        return v

    @staticmethod
    def barrier_1d(xmin, xmax, n, v0, x_lo, x_hi):
        v = np.empty(n)
        ref().ref_barrier_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(v0),
                             C.c_double(x_lo), C.c_double(x_hi), _p(v))
        return v

    @staticmethod
    def barrier_3d(dims, lo, hi, v0, x_lo, x_hi):
        v = np.empty(int(np.prod(dims)))
        ref().ref_barrier_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(v0),
                             C.c_double(x_lo), C.c_double(x_hi), _p(v))
        return v

    @staticmethod
    def regularized_coulomb_3d(dims, lo, hi, z, c):
        v = np.empty(int(np.prod(dims)))
        ref().ref_regularized_coulomb_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(z),
                                         _dvec(c), _p(v))
        return v

    @staticmethod
    def absorbing_mask_1d(xmin, xmax, n, width):
        m = np.empty(n)
        ref().ref_absorbing_mask_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(width),
                                    _p(m))
        return m

    @staticmethod
    def absorbing_mask_3d(dims, lo, hi, width):
        m = np.empty(int(np.prod(dims)))
        ref().ref_absorbing_mask_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(width), _p(m))
        return m

    @staticmethod
    def gaussian_1d(xmin, xmax, n, x0, sigma, k0):
        psi = np.empty(n, np.complex128)
        ref().ref_gaussian_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(x0),
                              C.c_double(sigma), C.c_double(k0), _p(psi))
        return psi

    @staticmethod
    def gaussian_3d(dims, lo, hi, r0, sigma, k0):
        psi = np.empty(int(np.prod(dims)), np.complex128)
        ref().ref_gaussian_3d(_ivec(dims), _dvec(lo), _dvec(hi), _dvec(r0), _dvec(sigma),
                              _dvec(k0), _p(psi))
        return psi

    @staticmethod
    def norm_sq_1d(psi, xmin, xmax):
        psi = _c128(psi)
        return ref().ref_norm_sq_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax))

    @staticmethod
    def norm_sq_3d(psi, dims, lo, hi):
        psi = _c128(psi)
        return ref().ref_norm_sq_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi))

    @staticmethod
    def mean_position_1d(psi, xmin, xmax):
        psi = _c128(psi)
        return ref().ref_mean_position_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax))

    @staticmethod
    def mean_energy_1d(psi, xmin, xmax, v):
        psi = _c128(psi)
        v = np.ascontiguousarray(v, dtype=np.float64)
        return ref().ref_mean_energy_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax),
                                        _p(v))

    @staticmethod
    def energy_variance_1d(psi, xmin, xmax, v):
        psi = _c128(psi)
        v = np.ascontiguousarray(v, dtype=np.float64)
        return ref().ref_energy_variance_1d(_p(psi), psi.size, C.c_double(xmin),
                                            C.c_double(xmax), _p(v))

    @staticmethod
    def probability_in_range_1d(psi, xmin, xmax, a, b):
        psi = _c128(psi)
        return ref().ref_probability_in_range_1d(_p(psi), psi.size, C.c_double(xmin),
                                                 C.c_double(xmax), C.c_double(a), C.c_double(b))

    @staticmethod
    def mean_position_3d(psi, dims, lo, hi):
        psi = _c128(psi)
        out = np.empty(3)
        ref().ref_mean_position_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(out))
        return out

    @staticmethod
    def mean_energy_3d(psi, dims, lo, hi, v, mass=1.0):
        psi = _c128(psi)
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        return ref().ref_mean_energy_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(v),
                                        C.c_double(mass))

    @staticmethod
    def it1d_relax(v, xmin, xmax, dtau, psi, nsteps):
        v = np.ascontiguousarray(v, dtype=np.float64)
        psi = _c128(psi).copy()
        if ref().ref_it1d_relax(_p(v), v.size, C.c_double(xmin), C.c_double(xmax),
                                C.c_double(dtau), _p(psi), nsteps) != 0:
            raise ValueError("size must be a power of two")
        return psi

    @staticmethod
    def it3d_tables(v, dims, lo, hi, dtau, mass=1.0):
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        hw = np.empty(v.size)
        kw = np.empty(v.size)
        ref().ref_it3d_tables(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(dtau),
                              C.c_double(mass), _p(hw), _p(kw))
        return hw, kw

    @staticmethod
    def it3d_relax(v, dims, lo, hi, dtau, psi, nsteps, mass=1.0, lower=()):
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        psi = _c128(psi).copy()
        shape = psi.shape
        psi = psi.ravel()
        low = [_c128(q).ravel() for q in lower]
        arr = (C.c_void_p * max(1, len(low)))(*[q.ctypes.data for q in low])
        if ref().ref_it3d_relax(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(dtau),
                                C.c_double(mass), _p(psi), arr, len(low), nsteps) != 0:
            raise ValueError("size must be a power of two")
        return psi.reshape(shape)

    @staticmethod
    def inner_product(a, b, dims, lo, hi):
        a = _c128(a).ravel()
        b = _c128(b).ravel()
        out = np.empty(2)
        ref().ref_inner_product(_p(a), _p(b), _ivec(dims), _dvec(lo), _dvec(hi), _p(out))
        return complex(out[0], out[1])

    @staticmethod
    def soft_coulomb_1d(xmin, xmax, n, z, a, x0=0.0):
        v = np.empty(n)
        ref().ref_soft_coulomb_1d(C.c_double(xmin), C.c_double(xmax), n, C.c_double(z),
                                  C.c_double(a), C.c_double(x0), _p(v))
        return v

    @staticmethod
    def soft_coulomb_3d(dims, lo, hi, z, a, c):
        v = np.empty(int(np.prod(dims)))
        ref().ref_soft_coulomb_3d(_ivec(dims), _dvec(lo), _dvec(hi), C.c_double(z), C.c_double(a),
                                  _dvec(c), _p(v))
        return v

    @staticmethod
    def sigma_x_1d(psi, xmin, xmax):
        psi = _c128(psi)
        return ref().ref_sigma_x_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax))

    @staticmethod
    def mean_momentum_1d(psi, xmin, xmax):
        psi = _c128(psi)
        return ref().ref_mean_momentum_1d(_p(psi), psi.size, C.c_double(xmin), C.c_double(xmax))

    @staticmethod
    def sigma_position_3d(psi, dims, lo, hi):
        psi = _c128(psi).ravel()
        out = np.empty(3)
        ref().ref_sigma_position_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(out))
        return tuple(out)

    @staticmethod
    def mean_momentum_3d(psi, dims, lo, hi):
        psi = _c128(psi).ravel()
        out = np.empty(3)
        ref().ref_mean_momentum_3d(_p(psi), _ivec(dims), _dvec(lo), _dvec(hi), _p(out))
        return tuple(out)

    @staticmethod
    def driven_evolve(v, dims, lo, hi, dt, psi, nsteps, axis, amplitude, omega, t0=0.0,
                      mass=1.0):
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        psi = _c128(psi).copy()
        shape = psi.shape
        psi = psi.ravel()
        if ref().ref_driven_step(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(dt),
                                 C.c_double(mass), _p(psi), _dvec(axis), C.c_double(amplitude),
                                 C.c_double(omega), C.c_double(t0), nsteps) != 0:
            raise ValueError("size must be a power of two")
        return psi.reshape(shape)

    @staticmethod
    def rotate_axis(psi, dims, lo, hi, axis, theta):
        psi = _c128(psi).copy()
        if ref().ref_rotate_axis(_p(psi.ravel()), _ivec(dims), _dvec(lo), _dvec(hi), axis,
                                 C.c_double(theta)) != 0:
            raise ValueError("size must be a power of two")
        return psi

    @staticmethod
    def magnetic_evolve(v, dims, lo, hi, dt, bfield, axis, psi, nsteps):
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        psi = _c128(psi).copy()
        shape = psi.shape
        psi = psi.ravel()
        veff = np.empty_like(v)
        if ref().ref_magnetic_evolve(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(dt),
                                     C.c_double(bfield), axis, _p(psi), nsteps, _p(veff)) != 0:
            raise ValueError("size must be a power of two")
        return psi.reshape(shape), veff

    @staticmethod
    def lat2d_evolve(v, dims, lo, hi, dt, psi, nsteps, relax=False, links=0, a=0.0, xs=0.0,
                     ys=0.0, cut_up=True):
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        psi = _c128(psi).copy()
        shape = psi.shape
        psi = psi.ravel()
        lx = np.empty(dims[0] * dims[1], np.complex128)
        e = C.c_double()
        ref().ref_lat2d_evolve(_p(v), _ivec(dims), _dvec(lo), _dvec(hi), C.c_double(dt), _p(psi),
                               nsteps, 1 if relax else 0, links, C.c_double(a), C.c_double(xs),
                               C.c_double(ys), 1 if cut_up else 0, _p(lx), C.byref(e))
        return psi.reshape(shape), e.value, lx
