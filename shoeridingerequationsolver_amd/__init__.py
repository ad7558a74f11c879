"""MI355X split-operator TDSE propagator: host-side mirror of the reference
core interface over the C ABI (include/tdse.h, libtdse.so).

Reference interface mirrored here (names, argument meaning, error cases):

* ``Grid1D`` / ``Grid3D``                  core/src/grid.ixx:10-30
* ``wavenumbers``                          core/src/spectral.ixx:14-23
* ``SplitOperator1D(grid, V, dt)``         core/src/propagator.ixx:21-64
* ``SplitOperator3D(grid, V, dt, mass)``   core/src/propagator.ixx:66-134
* ``WavepacketSimulation``                 core/src/simulation.ixx:22-88
* ``ImaginaryTimePropagator1D/3D``         core/src/imaginary_time.ixx:21-144
* ``DipoleDrive``, ``driven_step``         core/src/drive.ixx:19-66
* ``MagneticPropagator3D``                 core/src/magnetic.ixx:19-80
* ``PeierlsLattice2D``                     core/src/lattice2d.ixx:27-275
* potential / mask / packet builders       core/src/potential.ixx, wavepacket.ixx

Fields are numpy ``complex128`` arrays: shape ``(n,)`` for a ``Grid1D`` and
``(nz, ny, nx)`` for a ``Grid3D`` (x fastest, core/src/field.ixx).  ``step``
updates the array in place, like ``SplitOperator*::step(Field&, int)``.
Unsupported grid lengths raise ``ValueError`` (the reference raises
``std::invalid_argument`` for non powers of two, core/src/fft.ixx:42-44).

``Propagator`` is the device-resident form: the state stays in HBM between
calls, which is what the serving and batching layers use.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .capi import (TDSE_FP32, TDSE_FP64, TDSE_OBS_ENERGY, TDSE_OBS_MOMENTUM, TDSE_OBS_NORM,
                   TDSE_OBS_POSITION, TDSE_OBS_VARIANCE, TdseError)

__all__ = [
    "Grid1D", "Grid3D", "wavenumbers", "Context", "Propagator", "SplitOperator1D",
    "SplitOperator3D", "WavepacketSimulation", "harmonic_potential", "barrier_potential",
    "regularized_coulomb_potential", "coulomb_floor_potential", "absorbing_mask",
    "gaussian_wavepacket", "TDSE_FP32", "TDSE_FP64", "TdseError", "Batch",
    "ImaginaryTimePropagator1D", "ImaginaryTimePropagator3D", "soft_coulomb_potential",
    "MagneticPropagator3D", "PeierlsLattice2D", "DipoleDrive", "driven_step", "phase_tables",
]

PRECISIONS = {"fp32": TDSE_FP32, "fp64": TDSE_FP64, TDSE_FP32: TDSE_FP32, TDSE_FP64: TDSE_FP64}


# ---- grids (core/src/grid.ixx) ----------------------------------------------

@dataclass(frozen=True)
class Grid1D:
    xmin: float
    xmax: float
    n: int

    def size(self) -> int:
        return self.n

    def spacing(self) -> float:
        return (self.xmax - self.xmin) / self.n

    def coord(self, i):
        return self.xmin + i * self.spacing()


@dataclass(frozen=True)
class Grid3D:
    x: Grid1D
    y: Grid1D
    z: Grid1D

    def size(self) -> int:
        return self.x.n * self.y.n * self.z.n

    def flat(self, i, j, k):
        return i + self.x.n * (j + self.y.n * k)

    def cell_volume(self) -> float:
        return self.x.spacing() * self.y.spacing() * self.z.spacing()

    @property
    def shape(self):
        return (self.z.n, self.y.n, self.x.n)


def _axes(grid):
    """(n[3], lo[3], hi[3]) of a Grid1D (as n x 1 x 1) or Grid3D."""
    if isinstance(grid, Grid1D):
        return (grid.n, 1, 1), (grid.xmin, 0.0, 0.0), (grid.xmax, 1.0, 1.0)
    return ((grid.x.n, grid.y.n, grid.z.n), (grid.x.xmin, grid.y.xmin, grid.z.xmin),
            (grid.x.xmax, grid.y.xmax, grid.z.xmax))


def _shape(grid):
    return (grid.n,) if isinstance(grid, Grid1D) else grid.shape


def wavenumbers(g: Grid1D) -> np.ndarray:
    """fftfreq-ordered wavenumbers 2 pi / L * (j or j - n) (core/src/spectral.ixx:14-23)."""
    dk = 2.0 * math.pi / (g.xmax - g.xmin)
    j = np.arange(g.n)
    return dk * np.where(2 * j < g.n, j, j - g.n).astype(np.float64)


# ---- device context and resident propagator ---------------------------------

class Context:
    """A device and an optional HBM budget shared by its propagators."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        L = capi.lib()
        h = C.c_void_p()
        capi.check(L.tdse_ctx_create(device, hbm_budget_bytes, C.byref(h)))
        self._h = h
        self.device = device
        self.budget = hbm_budget_bytes

    @property
    def handle(self):
        return self._h

    def bytes_in_use(self) -> int:
        v = C.c_uint64()
        capi.check(capi.lib().tdse_ctx_bytes_in_use(self._h, C.byref(v)))
        return v.value

    def submit(self, props, nsteps):
        """Queue nsteps[i] steps on props[i] (each on its own stream; all
        arguments are checked first).  ``tdse_batch_submit``."""
        props, nsteps = list(props), [int(k) for k in nsteps]
        if len(props) != len(nsteps):
            raise ValueError("one step count per propagator")
        hs = (C.c_void_p * len(props))(*[p.handle for p in props])
        ks = (C.c_int * len(nsteps))(*nsteps)
        try:
            capi.check(capi.lib().tdse_batch_submit(self._h, hs, ks, len(props)))
        except TdseError as e:
            raise _map_error(e) from None

    def sync(self, timeout_ms: int = -1):
        """Wait for every plan and batch of the context (``tdse_sync``);
        TdseError with TDSE_E_TIMEOUT when work is still queued after
        timeout_ms (0: a poll)."""
        capi.check(capi.lib().tdse_sync(self._h, int(timeout_ms)))

    def close(self):
        if self._h:
            capi.lib().tdse_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT_CTX = None


def default_context() -> Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


def plan_bytes(grid, precision="fp32", with_mask=False) -> int:
    n, _, _ = _axes(grid)
    v = C.c_uint64()
    capi.check(capi.lib().tdse_plan_bytes(capi.ivec(n), PRECISIONS[precision], int(with_mask),
                                          C.byref(v)))
    return v.value


def _map_error(e: TdseError):
    if e.status in (capi.TDSE_E_UNSUPPORTED, capi.TDSE_E_INVALID):
        return ValueError(str(e))
    return e


class Propagator:
    """Device-resident split-operator propagator for one grid (one C-ABI plan)."""

    def __init__(self, grid, potential, dt, mass=1.0, precision="fp64", mask=None,
                 context: Context | None = None, unfused_baseline: bool = False,
                 imaginary_time: bool = False, bfield: float = 0.0, b_axis: int = 2,
                 lattice: bool = False):
        self.grid = grid
        self.ctx = context or default_context()
        self.precision = PRECISIONS[precision]
        self.dt = float(dt)
        self.mass = float(mass)
        n, lo, hi = _axes(grid)
        self.shape = _shape(grid)
        self.size = int(np.prod(n))
        L = capi.lib()
        h = C.c_void_p()
        try:
            capi.check(L.tdse_plan_create(self.ctx.handle, capi.ivec(n), capi.dvec(lo),
                                          capi.dvec(hi), C.c_double(dt), C.c_double(mass),
                                          self.precision,
                                          (capi.TDSE_FLAG_UNFUSED_BASELINE if unfused_baseline else 0)
                                          | (capi.TDSE_FLAG_IMAGINARY_TIME if imaginary_time else 0)
                                          | (capi.TDSE_FLAG_LATTICE if lattice else 0),
                                          C.byref(h)))
        except TdseError as e:
            raise _map_error(e) from None
        self._h = h
        self._has_mask = False
        if bfield != 0.0:
            try:
                capi.check(L.tdse_plan_set_magnetic(self._h, C.c_double(bfield), int(b_axis)))
            except TdseError as e:
                raise _map_error(e) from None
        if potential is not None:
            self.set_potential(potential)
        if mask is not None:
            self.set_mask(mask)

    @property
    def handle(self):
        return self._h

    def _vec(self, a, name):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != self.size:
            raise ValueError(f"{name} has {a.size} entries, grid has {self.size}")
        return a

    def set_potential(self, v):
        v = self._vec(v, "potential")
        capi.check(capi.lib().tdse_plan_set_potential(self._h, capi.dptr(v)))

    def build(self, recipe):
        """Build potential, mask and the normalised Gaussian state on the GPU
        from a ``capi.Recipe`` (see ``workloads.Solve.recipe``)."""
        capi.check(capi.lib().tdse_plan_build(self._h, C.byref(recipe)))
        self._has_mask = self._has_mask or recipe.mask_width > 0.0

    def set_mask(self, m):
        if m is None:
            capi.check(capi.lib().tdse_plan_set_mask(self._h, None))
            self._has_mask = False
            return
        m = self._vec(m, "mask")
        capi.check(capi.lib().tdse_plan_set_mask(self._h, capi.dptr(m)))
        self._has_mask = True

    def upload(self, psi):
        a = np.ascontiguousarray(psi, dtype=np.complex128).ravel()
        if a.size != self.size:
            raise ValueError(f"state has {a.size} points, grid has {self.size}")
        capi.check(capi.lib().tdse_state_upload(self._h, capi.dptr(a)))

    def download(self, out=None):
        if out is None:
            out = np.empty(self.shape, np.complex128)
        if not (out.flags.c_contiguous and out.dtype == np.complex128 and out.size == self.size):
            raise ValueError("output must be a contiguous complex128 array of the grid's size")
        capi.check(capi.lib().tdse_state_download(self._h, capi.dptr(out)))
        return out

    def download_f32(self) -> np.ndarray:
        """The state as complex64 (``tdse_state_download_f32``): as stored for
        FP32 plans, rounded to nearest for FP64 plans."""
        out = np.empty(self.shape, np.complex64)
        capi.check(capi.lib().tdse_state_download_f32(
            self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def step(self, nsteps=1):
        capi.check(capi.lib().tdse_step(self._h, int(nsteps)))

    def copy_state_from(self, other: "Propagator", kernel: bool = False):
        """This plan's state = ``other``'s (same shape and precision), on the GPU
        (``kernel``: by the library's streaming copy kernel)."""
        fn = capi.lib().tdse_state_copy_kernel if kernel else capi.lib().tdse_state_copy
        try:
            capi.check(fn(self._h, other.handle))
        except TdseError as e:
            raise _map_error(e) from None

    def sync(self, timeout_ms=-1):
        capi.check(capi.lib().tdse_plan_sync(self._h, int(timeout_ms)))

    def done(self) -> bool:
        """True once the plan's queued work has finished (non-blocking)."""
        st = capi.lib().tdse_plan_sync(self._h, 0)
        if st == capi.TDSE_E_TIMEOUT:
            return False
        capi.check(st)
        return True

    def scale(self, s):
        capi.check(capi.lib().tdse_state_scale(self._h, C.c_double(s)))

    def observe(self, what=TDSE_OBS_NORM | TDSE_OBS_POSITION | TDSE_OBS_ENERGY):
        o = capi.Observables()
        capi.check(capi.lib().tdse_observe(self._h, what, C.byref(o)))
        return {"norm": o.norm, "sum_abs2": o.sum_abs2, "position": tuple(o.position),
                "potential": o.potential, "kinetic": o.kinetic, "energy": o.energy,
                "sigma": tuple(o.sigma), "momentum": tuple(o.momentum),
                "energy_variance": o.energy_variance, "peak": o.peak}

    def probability_in_slab(self, axis, lo, hi):
        v = C.c_double()
        capi.check(capi.lib().tdse_probability_in_slab(self._h, axis, lo, hi, C.byref(v)))
        return v.value

    def probability_in_box(self, lo, hi):
        """Probability lo[a] <= x_a < hi[a] on every axis of extent above 1,
        times the cell volume (``tdse_probability_in_box``)."""
        v = C.c_double()
        capi.check(capi.lib().tdse_probability_in_box(self._h, capi.dvec(lo), capi.dvec(hi),
                                                       C.byref(v)))
        return v.value

    def stream(self):
        s = C.c_void_p()
        capi.check(capi.lib().tdse_plan_stream(self._h, C.byref(s)))
        return s.value

    def device_ptr(self):
        s = C.c_void_p()
        capi.check(capi.lib().tdse_state_device_ptr(self._h, C.byref(s)))
        return s.value

    def step_bytes(self):
        v = C.c_uint64()
        capi.check(capi.lib().tdse_plan_step_bytes(self._h, C.byref(v)))
        return v.value

    def profile(self, nsteps):
        """Per-sweep event timing over nsteps: {slot name: (total ms, launches)}."""
        ms = (C.c_double * 16)()
        cnt = (C.c_int * 16)()
        ns = C.c_int()
        capi.check(capi.lib().tdse_profile_step(self._h, int(nsteps), 16, ms, cnt, C.byref(ns)))
        L = capi.lib()
        return {L.tdse_plan_sweep_name(self._h, i).decode(): (ms[i], cnt[i])
                for i in range(ns.value)}

    def set_drive(self, axis, amplitude, omega):
        """Dipole drive amplitude (axis . r) cos(omega t) on the following
        steps (core/src/drive.ixx); amplitude 0 removes it."""
        try:
            capi.check(capi.lib().tdse_plan_set_drive(self._h, capi.dvec(axis), C.c_double(amplitude),
                                                      C.c_double(omega)))
        except TdseError as e:
            raise _map_error(e) from None

    @property
    def time(self) -> float:
        v = C.c_double()
        capi.check(capi.lib().tdse_plan_get_time(self._h, C.byref(v)))
        return v.value

    @time.setter
    def time(self, t):
        capi.check(capi.lib().tdse_plan_set_time(self._h, C.c_double(t)))

    def set_uniform_field(self, b):
        """Lattice plans: uniform field along z in the Landau gauge."""
        try:
            capi.check(capi.lib().tdse_lattice_set_uniform_field(self._h, C.c_double(b)))
        except TdseError as e:
            raise _map_error(e) from None

    def set_solenoid(self, phi, xs, ys, cut_up=True):
        """Lattice plans: a solenoid of flux phi at (xs, ys), string gauge."""
        try:
            capi.check(capi.lib().tdse_lattice_set_solenoid(self._h, C.c_double(phi), C.c_double(xs),
                                                            C.c_double(ys), 1 if cut_up else 0))
        except TdseError as e:
            raise _map_error(e) from None

    def relax_stats(self) -> dict:
        """Imaginary time: the last step's norm before renormalisation, the
        energy estimate -ln(norm)/(2 dtau) and the peak density."""
        st = capi.RelaxStats()
        try:
            capi.check(capi.lib().tdse_plan_relax_stats(self._h, C.byref(st)))
        except TdseError as e:
            raise _map_error(e) from None
        return {"norm_sq": st.norm_sq, "energy": st.energy, "peak": st.peak}

    def relax_deflated(self, lower, nsteps):
        """Imaginary time: relax with Gram-Schmidt deflation against the
        current states of the ``lower`` propagators (same shape/precision)."""
        arr = (C.c_void_p * max(1, len(lower)))(*[q.handle for q in lower])
        try:
            capi.check(capi.lib().tdse_relax_deflated(self._h, arr, len(lower), int(nsteps)))
        except TdseError as e:
            raise _map_error(e) from None

    @property
    def variants(self) -> dict:
        """Kernel variant per schedule slot: {slot name: variant}."""
        v = (C.c_int * 16)()
        ns = C.c_int()
        L = capi.lib()
        capi.check(L.tdse_plan_get_variants(self._h, 16, v, C.byref(ns)))
        return {L.tdse_plan_sweep_name(self._h, i).decode(): v[i] for i in range(ns.value)}

    def set_variants(self, variants: dict):
        """Select kernel variants by slot name (every slot of the schedule)."""
        L = capi.lib()
        names = list(self.variants)
        if set(variants) != set(names):
            raise ValueError(f"expected variants for slots {names}")
        arr = (C.c_int * len(names))(*[int(variants[k]) for k in names])
        try:
            capi.check(L.tdse_plan_set_variants(self._h, arr, len(names)))
        except TdseError as e:
            raise _map_error(e) from None

    def tuning_key(self) -> str:
        """Key of this plan's sweep shapes in a tuning cache."""
        prec = "fp64" if self.precision == capi.TDSE_FP64 else "fp32"
        n, _, _ = _axes(self.grid)
        return f"{prec}:{n[0]}x{n[1]}x{n[2]}:{'mask' if self._has_mask else 'nomask'}"

    def autotune(self, reps: int = 6, cache: str | None = None) -> dict:
        """Time every compiled kernel variant of each sweep on a scratch copy of
        the state and keep the fastest.  With `cache` (a JSON file), a stored
        choice for the same precision, shape and mask presence is applied
        without timing, and a new choice is stored."""
        import json
        import os
        key = self.tuning_key()
        table = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                table = json.load(f)
            if key in table:
                self.set_variants(table[key])
                return dict(table[key])
        v = (C.c_int * 16)()
        ns = C.c_int()
        L = capi.lib()
        capi.check(L.tdse_plan_autotune(self._h, int(reps), 16, v, C.byref(ns)))
        got = {L.tdse_plan_sweep_name(self._h, i).decode(): v[i] for i in range(ns.value)}
        if cache:
            table[key] = got
            tmp = cache + ".tmp"
            with open(tmp, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return got

    def close(self):
        if getattr(self, "_h", None):
            capi.lib().tdse_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many device-resident propagators advanced together (one C-ABI batch).

    ``step(n)`` advances every member by n Strang steps, exactly as
    ``Propagator.step(n)`` would one by one; members sharing a sweep shape are
    swept by one batched launch.  Captured-graph replay is switched off."""

    def __init__(self, props, context: Context | None = None):
        self.props = list(props)
        self.ctx = context or self.props[0].ctx
        arr = (C.c_void_p * len(self.props))(*[p.handle for p in self.props])
        h = C.c_void_p()
        try:
            capi.check(capi.lib().tdse_batch_create(self.ctx.handle, arr, len(self.props),
                                                    C.byref(h)))
        except TdseError as e:
            raise _map_error(e) from None
        self._h = h

    def step(self, nsteps, use_graph=False):
        capi.check(capi.lib().tdse_batch_step(self._h, int(nsteps), int(bool(use_graph))))

    def sync(self, timeout_ms=-1):
        capi.check(capi.lib().tdse_batch_sync(self._h, int(timeout_ms)))

    def stream(self):
        s = C.c_void_p()
        capi.check(capi.lib().tdse_batch_stream(self._h, C.byref(s)))
        return s.value

    def launches_per_round(self):
        v = C.c_int()
        capi.check(capi.lib().tdse_batch_launches_per_round(self._h, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "_h", None):
            capi.lib().tdse_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- drop-in propagators (core/src/propagator.ixx) --------------------------

def phase_tables(grid, potential, dt, mass=1.0):
    """The reference propagator's FP64 tables (core/src/propagator.ixx:23-40,
    72-101): (half_v, kinetic) as complex128 arrays in the field's shape,
    bit-identical to ``half_potential_phase()`` / ``kinetic_phase()``."""
    n, lo, hi = _axes(grid)
    v = np.ascontiguousarray(potential, dtype=np.float64).ravel()
    size = int(np.prod(n))
    if v.size != size:
        raise ValueError(f"potential has {v.size} entries, grid has {size}")
    half = np.empty(size, np.complex128)
    kin = np.empty(size, np.complex128)
    capi.check(capi.lib().tdse_build_phase_tables(
        capi.ivec(n), capi.dvec(lo), capi.dvec(hi), capi.dptr(v), C.c_double(dt),
        C.c_double(mass), capi.dptr(half.view(np.float64)), capi.dptr(kin.view(np.float64))))
    return half.reshape(_shape(grid)), kin.reshape(_shape(grid))


class _SplitOperatorBase:
    def __init__(self, grid, potential, dt, mass, precision, context):
        self._grid = grid
        self._dt = float(dt)
        self._mass = float(mass)
        self._v = np.ascontiguousarray(potential, dtype=np.float64).ravel()
        self._tables = None
        self._prop = Propagator(grid, self._v, dt, mass=mass, precision=precision,
                                context=context)

    def dt(self) -> float:
        return self._dt

    def device(self) -> Propagator:
        """The device plan behind this propagator."""
        return self._prop

    def _phase_tables(self):
        if self._tables is None:
            self._tables = phase_tables(self._grid, self._v, self._dt, self._mass)
        return self._tables

    def half_potential_phase(self) -> np.ndarray:
        """cis(-0.5 V dt) per cell, FP64 (core/src/propagator.ixx:107); built on
        first use on the host (the GPU path keeps reduced kick angles)."""
        return self._phase_tables()[0]

    def kinetic_phase(self) -> np.ndarray:
        """cis(-0.5 |k|^2 / m dt) over the full grid, FP64 (core/src/propagator.ixx:108)."""
        return self._phase_tables()[1]

    def step(self, psi: np.ndarray, nsteps: int = 1) -> None:
        """Advance ``psi`` in place by ``nsteps`` Strang steps."""
        if nsteps < 0:
            raise ValueError("nsteps must be >= 0")
        if psi.dtype != np.complex128 or not psi.flags.c_contiguous:
            raise ValueError("psi must be a contiguous complex128 array")
        if psi.size != self._prop.size:
            raise ValueError("psi does not match the grid")
        if nsteps == 0:
            return
        self._prop.upload(psi)
        self._prop.step(nsteps)
        self._prop.download(psi.reshape(self._prop.shape))


class SplitOperator1D(_SplitOperatorBase):
    """``ses::SplitOperator1D(grid, V, dt)`` (core/src/propagator.ixx:21-64)."""

    def __init__(self, grid: Grid1D, potential, dt, precision="fp64", context=None):
        super().__init__(grid, potential, dt, 1.0, precision, context)


class SplitOperator3D(_SplitOperatorBase):
    """``ses::SplitOperator3D(grid, V, dt, mass)`` (core/src/propagator.ixx:66-134)."""

    def __init__(self, grid: Grid3D, potential, dt, mass=1.0, precision="fp64", context=None):
        super().__init__(grid, potential, dt, mass, precision, context)


@dataclass
class DipoleDrive:
    """``ses::DipoleDrive`` (core/src/drive.ixx:19-23): V = amplitude (axis . r)
    cos(omega t); the axis need not be a unit vector (it scales E0)."""
    axis: tuple = (0.0, 0.0, 1.0)
    amplitude: float = 0.0
    omega: float = 0.0


def driven_step(psi: np.ndarray, prop: SplitOperator3D, d: DipoleDrive, t0: float,
                nsteps: int) -> None:
    """``ses::driven_step(psi, prop, d, t0, nsteps)`` (core/src/drive.ixx:52-66):
    ``nsteps`` Strang steps of ``prop`` from time ``t0`` with the dipole
    half-kicks around each, in place.  The drive's kicks ride in the kick
    sweeps; ``prop`` is left without a drive afterwards."""
    if nsteps < 0:
        raise ValueError("nsteps must be >= 0")
    if nsteps == 0:
        return
    p = prop._prop
    p.set_drive(tuple(d.axis), d.amplitude, d.omega)
    try:
        p.time = t0
        prop.step(psi, nsteps)
    finally:
        p.set_drive(tuple(d.axis), 0.0, 0.0)


class PeierlsLattice2D:
    """``ses::PeierlsLattice2D(grid, V, dt)`` (core/src/lattice2d.ixx:31-275): 2D
    finite differences with Peierls link phases on an n x m x 1 grid, open
    boundaries.  ``step`` propagates, ``relax`` relaxes in imaginary time (a
    second plan with the same links), ``energy`` is the lattice <H>."""

    def __init__(self, grid: Grid3D, potential, dt, precision="fp64", context=None):
        self._grid = grid
        self._dt = float(dt)
        self._v = np.ascontiguousarray(potential, dtype=np.float64).ravel()
        self._prec = precision
        self._ctx = context
        self._links = ("none",)
        self._prop = Propagator(grid, self._v, dt, precision=precision, context=context,
                                lattice=True)
        self._relaxer = None

    def dt(self) -> float:
        return self._dt

    def _apply_links(self, p):
        kind = self._links[0]
        if kind == "uniform":
            p.set_uniform_field(self._links[1])
        elif kind == "solenoid":
            p.set_solenoid(*self._links[1:])
        else:
            p.set_solenoid(0.0, 1e300, 1e300)  # outside the lattice: every link 1

    def set_solenoid(self, phi, xs, ys, cut_up=True):
        self._links = ("solenoid", phi, xs, ys, cut_up)
        self._apply_links(self._prop)
        if self._relaxer:
            self._apply_links(self._relaxer)

    def set_uniform_field(self, b):
        self._links = ("uniform", b)
        self._apply_links(self._prop)
        if self._relaxer:
            self._apply_links(self._relaxer)

    def link_x(self, i, j) -> complex:
        """Directed x link (i, j) -> (i + 1, j) (host copy of the plan's rule)."""
        kind = self._links[0]
        if kind == "uniform":
            th = -self._links[1] * self._grid.x.spacing() * self._grid.y.coord(j)
            return complex(math.cos(th), math.sin(th))
        if kind == "solenoid":
            phi, xs, ys, cut_up = self._links[1:]
            gx, gy = self._grid.x, self._grid.y
            isol = max([i_ for i_ in range(gx.n - 1) if gx.coord(i_) <= xs < gx.coord(i_ + 1)],
                       default=-1)
            jsol = max([j_ for j_ in range(gy.n - 1) if gy.coord(j_) <= ys < gy.coord(j_ + 1)],
                       default=-1)
            if isol >= 0 and jsol >= 0 and i == isol and ((cut_up and j > jsol) or
                                                          (not cut_up and j <= jsol)):
                return complex(math.cos(phi), -math.sin(phi) if cut_up else math.sin(phi))
        return 1.0 + 0.0j

    def link_y(self, i, j) -> complex:
        return 1.0 + 0.0j

    def _run(self, p, psi, nsteps):
        if psi.dtype != np.complex128 or not psi.flags.c_contiguous or psi.size != p.size:
            raise ValueError("psi must be a contiguous complex128 array of the grid's size")
        if nsteps <= 0:
            return
        p.upload(psi)
        p.step(nsteps)
        p.download(psi.reshape(p.shape))

    def step(self, psi: np.ndarray, nsteps: int = 1) -> None:
        self._run(self._prop, psi, nsteps)

    def relax(self, psi: np.ndarray, nsteps: int = 1) -> None:
        if self._relaxer is None:
            self._relaxer = Propagator(self._grid, self._v, self._dt, precision=self._prec,
                                       context=self._ctx, lattice=True, imaginary_time=True)
            self._apply_links(self._relaxer)
        self._run(self._relaxer, psi, nsteps)

    def energy(self, psi: np.ndarray) -> float:
        self._prop.upload(np.ascontiguousarray(psi, dtype=np.complex128))
        return self._prop.observe(TDSE_OBS_NORM | TDSE_OBS_ENERGY)["energy"]


class MagneticPropagator3D:
    """``ses::MagneticPropagator3D(grid, V, dt, B, axis)`` (core/src/magnetic.ixx:19-80):
    uniform field along ``axis`` in the symmetric gauge; each step is
    R(a) . Strang(V + B^2/8 r_perp^2) . R(a) with an exact three-shear rotation."""

    def __init__(self, grid: Grid3D, potential, dt, bfield, axis=2, precision="fp64",
                 context=None):
        self._grid = grid
        self._dt = float(dt)
        self._v = np.ascontiguousarray(potential, dtype=np.float64).ravel()
        self._b = float(bfield)
        self._axis = int(axis)
        self._prop = Propagator(grid, self._v, dt, precision=precision, context=context,
                                bfield=bfield, b_axis=axis)

    def dt(self) -> float:
        return self._dt

    def effective_potential(self) -> np.ndarray:
        """V + B^2/8 |r_perp|^2 (host copy of what the plan holds)."""
        if self._b == 0.0:
            return self._v.copy()
        axes = [self._grid.x, self._grid.y, self._grid.z]
        co = [a.xmin + np.arange(a.n) * a.spacing() for a in axes]
        z, y, x = np.meshgrid(co[2], co[1], co[0], indexing="ij")
        c = (x, y, z)
        perp2 = sum(c[a] ** 2 for a in range(3) if a != self._axis)
        return self._v + (self._b * self._b / 8.0 * perp2).ravel()

    def step(self, psi: np.ndarray, nsteps: int = 1) -> None:
        if psi.dtype != np.complex128 or not psi.flags.c_contiguous or psi.size != self._prop.size:
            raise ValueError("psi must be a contiguous complex128 array of the grid's size")
        if nsteps <= 0:
            return
        self._prop.upload(psi)
        self._prop.step(nsteps)
        self._prop.download(psi.reshape(self._prop.shape))


# ---- imaginary time (core/src/imaginary_time.ixx) ----------------------------

class _ImaginaryTimeBase:
    def __init__(self, grid, potential, dtau, mass, precision, context):
        self._grid = grid
        self._dtau = float(dtau)
        self._mass = float(mass)
        self._v = np.ascontiguousarray(potential, dtype=np.float64).ravel()
        self._ctx = context
        self._prec = precision
        self._prop = Propagator(grid, self._v, dtau, mass=mass, precision=precision,
                                context=context, imaginary_time=True)

    def _check(self, psi, nsteps):
        if nsteps < 0:
            raise ValueError("nsteps must be >= 0")
        if psi.dtype != np.complex128 or not psi.flags.c_contiguous:
            raise ValueError("psi must be a contiguous complex128 array")
        if psi.size != self._prop.size:
            raise ValueError("psi does not match the grid")

    def half_potential_weight(self) -> np.ndarray:
        """exp(-V dtau / 2) per cell (host copy of the weight the GPU forms)."""
        return np.exp(-0.5 * self._v * self._dtau)

    def relax(self, psi: np.ndarray, nsteps: int) -> None:
        """``nsteps`` relax steps of ``psi`` in place, each ending normalised."""
        self._check(psi, nsteps)
        if nsteps == 0:
            return
        self._prop.upload(psi)
        self._prop.step(nsteps)
        self._prop.download(psi.reshape(self._prop.shape))

    def stats(self) -> dict:
        return self._prop.relax_stats()


class ImaginaryTimePropagator1D(_ImaginaryTimeBase):
    """``ses::ImaginaryTimePropagator1D(grid, V, dtau)`` (core/src/imaginary_time.ixx:21-63)."""

    def __init__(self, grid: Grid1D, potential, dtau, precision="fp64", context=None):
        super().__init__(grid, potential, dtau, 1.0, precision, context)


class ImaginaryTimePropagator3D(_ImaginaryTimeBase):
    """``ses::ImaginaryTimePropagator3D(grid, V, dtau, mass)``
    (core/src/imaginary_time.ixx:65-144)."""

    def __init__(self, grid: Grid3D, potential, dtau, mass=1.0, precision="fp64", context=None):
        super().__init__(grid, potential, dtau, mass, precision, context)

    def kinetic_weight(self) -> np.ndarray:
        """exp(-|k|^2 dtau / 2m) over the full grid, x fastest."""
        kx, ky, kz = (wavenumbers(a) for a in (self._grid.x, self._grid.y, self._grid.z))
        k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
        return np.exp(-0.5 * k2 / self._mass * self._dtau).ravel()

    def relax_deflated(self, psi: np.ndarray, lower, nsteps: int) -> None:
        """Relax with Gram-Schmidt deflation against the fields in ``lower``
        (core/src/imaginary_time.ixx:124-141): converges to the next state up."""
        self._check(psi, nsteps)
        holders = []
        try:
            for q in lower:
                h = Propagator(self._grid, None, self._dtau, precision=self._prec,
                               context=self._ctx)
                h.upload(np.ascontiguousarray(q, dtype=np.complex128))
                holders.append(h)
            self._prop.upload(psi)
            self._prop.relax_deflated(holders, nsteps)
            self._prop.download(psi.reshape(self._prop.shape))
        finally:
            for h in holders:
                h.close()


# ---- builders (core/src/potential.ixx, core/src/wavepacket.ixx) -------------

def _build(fn, grid, *scalars_and_vecs, out_complex=False):
    n, lo, hi = _axes(grid)
    out = np.empty(int(np.prod(n)) * (2 if out_complex else 1), np.float64)
    capi.check(fn(capi.ivec(n), capi.dvec(lo), capi.dvec(hi), *scalars_and_vecs, capi.dptr(out)))
    return out.view(np.complex128).reshape(_shape(grid)) if out_complex else out


def _center(grid, c):
    if isinstance(grid, Grid1D):
        return capi.dvec((float(c), 0.0, 0.0))
    return capi.dvec(c)


def harmonic_potential(grid, omega, c=0.0):
    """0.5 w^2 |r - c|^2 (core/src/potential.ixx:16-48); c is x0 for a Grid1D."""
    return _build(capi.lib().tdse_build_harmonic, grid, C.c_double(omega), _center(grid, c))


def barrier_potential(grid, v0, x_lo, x_hi):
    """v0 on x_lo <= x < x_hi along x (core/src/potential.ixx:102-132)."""
    return _build(capi.lib().tdse_build_barrier, grid, C.c_double(v0), C.c_double(x_lo),
                  C.c_double(x_hi))


def regularized_coulomb_potential(grid: Grid3D, Z, c):
    """-Z/r with the nucleus cell's analytic average (core/src/potential.ixx:134-154)."""
    return _build(capi.lib().tdse_build_regularized_coulomb, grid, C.c_double(Z), capi.dvec(c))


def coulomb_floor_potential(grid: Grid3D, Z, c):
    """-Z / max(r, h/2): the hydrogen-like workload's Coulomb (no reference builder)."""
    return _build(capi.lib().tdse_build_coulomb_floor, grid, C.c_double(Z), capi.dvec(c))


def soft_coulomb_potential(grid, Z, a, c=0.0):
    """-Z / sqrt(|r - c|^2 + a^2) (core/src/potential.ixx:25-33, 51-65); c is x0 for a
    Grid1D."""
    return _build(capi.lib().tdse_build_soft_coulomb, grid, C.c_double(Z), C.c_double(a),
                  _center(grid, c))


def absorbing_mask(grid, width):
    """sin^2 ramp to 0 within ``width`` of each wall (core/src/potential.ixx:194-233)."""
    return _build(capi.lib().tdse_build_absorbing_mask, grid, C.c_double(width))


def gaussian_wavepacket(grid, r0, sigma, k0):
    """Normalised Gaussian packet (core/src/wavepacket.ixx:16-48); scalars for a Grid1D."""
    if isinstance(grid, Grid1D):
        r0, sigma, k0 = (r0, 1.0, 1.0), (sigma, 1.0, 1.0), (k0, 0.0, 0.0)
    return _build(capi.lib().tdse_build_gaussian, grid, capi.dvec(r0), capi.dvec(sigma),
                  capi.dvec(k0), out_complex=True)


# ---- evolve wrapper (core/src/simulation.ixx) --------------------------------

@dataclass
class SimulationConfig:
    grid: Grid3D
    potential: np.ndarray
    r0: tuple
    sigma: tuple
    k0: tuple
    dt: float
    precision: str = "fp64"


class WavepacketSimulation:
    """``ses::WavepacketSimulation`` (core/src/simulation.ixx:22-88): owns the
    grid, potential, propagator and state; ``advance(n)`` steps it and
    ``relax(n, dtau)`` relaxes it in imaginary time.  The state stays on the
    GPU between calls and is read back by ``psi()``."""

    Config = SimulationConfig

    def __init__(self, cfg: SimulationConfig, context=None):
        self._grid = cfg.grid
        self._dt = float(cfg.dt)
        self._potential = np.ascontiguousarray(cfg.potential, dtype=np.float64)
        self._precision = cfg.precision
        self._ctx = context
        self._op = SplitOperator3D(cfg.grid, self._potential, cfg.dt, precision=cfg.precision,
                                   context=context)
        self._prop = self._op.device()
        self._prop.upload(gaussian_wavepacket(cfg.grid, cfg.r0, cfg.sigma, cfg.k0))
        self._steps = 0
        self._relaxer = None
        self._relaxer_dtau = 0.0

    def advance(self, nsteps: int) -> None:
        self._prop.step(nsteps)
        self._steps += nsteps

    def relax(self, nsteps: int, dtau: float) -> None:
        """``nsteps`` imaginary-time steps of the state (simulation.ixx:45-52):
        the relaxer is rebuilt when dtau changes; the clock does not move.
        The state moves between the two plans on the GPU."""
        if self._relaxer is None or self._relaxer_dtau != dtau:
            if self._relaxer is not None:
                self._relaxer._prop.close()
            self._relaxer = ImaginaryTimePropagator3D(self._grid, self._potential, dtau,
                                                      precision=self._precision,
                                                      context=self._ctx)
            self._relaxer_dtau = dtau
        if nsteps <= 0:
            return
        r = self._relaxer._prop
        r.copy_state_from(self._prop)
        r.step(nsteps)
        self._prop.copy_state_from(r)

    def time(self) -> float:
        return self._steps * self._dt

    def dt(self) -> float:
        return self._dt

    def grid(self):
        return self._grid

    def potential(self):
        return self._potential

    def psi(self) -> np.ndarray:
        return self._prop.download()

    def set_psi(self, psi) -> None:
        self._prop.upload(psi)

    def density(self) -> np.ndarray:
        return np.abs(self.psi()) ** 2

    def propagator(self) -> SplitOperator3D:
        """The simulation's SplitOperator3D (its phase tables, its device plan)."""
        return self._op
