"""In-run accumulated observables on the device (results/accumulators.py:12-188).

``result.accumulators`` maps a registered name to its parameters (``start``, ``stop``,
``every`` in ps, plus the accumulator's own keys); the controller samples each one on its
schedule between solver chunks (simulation_controller.py:1015-1057).  The sums live in
device memory in double precision and are updated by HIP kernels
(``pgpe_accumulate_power``, ``pgpe_accumulate_moments``; the momentum distribution's
transform by ``pgpe_fft2``; the spectra's line cuts, windowed blocks and power by
``pgpe_spectrum_*`` and ``pgpe_fft_columns``); only ``results()`` brings them to the host,
once, to normalise them.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native as N
from .device import DeviceArray, Stream
from .solver_api import Registry


@dataclass(frozen=True)
class AccumulatedResult:
    """One finalized accumulated array and its axis metadata (accumulators.py:12-21)."""
    name: str
    value: object
    dims: tuple
    units: str
    label: str = ""
    wigner_corrected: bool = False
    linear: bool = False


available_accumulators: Registry = Registry("accumulator")


def register_accumulator(name: str):
    """Register an accumulator class under a configuration name."""
    return available_accumulators.register(name)


def g1(n_k, *, ndim=2):
    """Normalised, centred first-order spatial correlation of a (shifted) momentum
    distribution (analysis/correlations.py:11-21)."""
    n_k = np.asarray(n_k)
    axes = tuple(range(-ndim, 0))
    corr = np.fft.ifftn(np.fft.ifftshift(n_k, axes=axes), axes=axes)
    origin = np.asarray(corr[(...,) + (0,) * ndim])
    den = np.where(np.abs(origin) > 1e-30, origin, 1)
    expand = (...,) + (None,) * ndim
    return np.fft.fftshift(np.where(np.abs(origin)[expand] > 1e-30, corr / den[expand], 0), axes=axes)


def g2(mean, squared_mean):
    """Second-order correlation, zero where the density is negligible (correlations.py:24-29)."""
    mean = np.asarray(mean)
    squared_mean = np.asarray(squared_mean)
    return np.where(mean > 1e-30, squared_mean / np.maximum(mean * mean, 1e-60), 0)


def _fft_length_ok(n) -> bool:
    """Transform lengths of the LDS FFT: 2^a (2..4096) or 3 * 2^a (6..3072)."""
    if isinstance(n, bool) or not isinstance(n, int) or n < 2:
        return False
    m = n // 3 if n % 3 == 0 else n
    return m & (m - 1) == 0 and n <= (3072 if n % 3 == 0 else 4096)


@dataclass
class SampleClock:
    """How many samples an accumulator holds and when the first and last were taken."""
    samples: int = 0
    t_first: float | None = None
    t_last: float | None = None

    def tick(self, t) -> None:
        self.t_first = t if self.samples == 0 else self.t_first
        self.t_last = t
        self.samples += 1

    def as_dict(self) -> dict:
        return {"samples": self.samples, "t_first": self.t_first, "t_last": self.t_last}

    def restore(self, values) -> None:
        self.samples, self.t_first, self.t_last = int(values["samples"]), values["t_first"], values["t_last"]


@dataclass
class FieldFrame:
    """Axes and constants of the accumulated field: public dims and shape, cell area and grid
    point count, and the per-member Wigner half-quantum (strength / cell, broadcast over the
    public shape) that the corrected results subtract."""
    dims: tuple
    shape: tuple
    points: int
    cell: float
    cells: int
    delta: np.ndarray
    strength: np.ndarray
    wigner: bool
    precision: int

    @classmethod
    def of(cls, controller) -> "FieldFrame":
        state, grid, cfg = controller.state, controller.grid, controller.cfg
        dims = ("member",) * state.batched + ("component",) * (state.components == 2) + ("y", "x")
        shape = state._public_shape
        wigner_on = lambda m: m.noise.enabled and m.noise.stochastic == "wigner"
        per = np.asarray([m.noise.stochastic_strength if wigner_on(m) else 0.0
                          for m in (controller._member_cfgs or [cfg])], dtype=np.float64)
        lead = len(per) if state.batched else 1
        cell = grid.dx * grid.dy
        strength = per.reshape((lead,) + (1,) * (len(shape) - 1)) if state.batched else \
            per.reshape((1,) * len(shape))
        return cls(dims=dims, shape=shape, points=int(np.prod(state._canonical_shape)), cell=cell,
                   cells=grid.nx * grid.ny, delta=strength / cell, strength=strength, wigner=bool(wigner_on(cfg)),
                   precision=0 if state.precision == "single" else 1)


class Accumulator:
    """A device-side observable sampled on the controller's start / every / stop schedule
    (results/accumulators.py:32-91 defines the contract: update, state_dict, results)."""
    name = ""
    parameter_keys: frozenset = frozenset()

    def __init__(self, controller, *, start=0.0, stop=None, every=None, **params):
        if params:
            raise ValueError(f"Unknown {self.name} parameters: {sorted(params)}")
        self.controller = controller
        self.start, self.stop, self.every = start, stop, every
        self.clock = SampleClock()
        self.frame = FieldFrame.of(controller)
        self.ndim = 2
        self.stream = Stream()

    # the schedule and result writers read these
    samples = property(lambda self: self.clock.samples)
    t_first = property(lambda self: self.clock.t_first)
    t_last = property(lambda self: self.clock.t_last)
    dims = property(lambda self: self.frame.dims)
    shape = property(lambda self: self.frame.shape)
    wigner_corrected = property(lambda self: self.frame.wigner)

    def update(self, state, t) -> None:
        raise NotImplementedError

    def device_sums(self) -> dict:
        """Name -> DeviceArray of the running sums a checkpoint stores."""
        return {}

    def state_dict(self) -> dict:
        return {**self.clock.as_dict(), **{k: v.get() for k, v in self.device_sums().items()}}

    def load_state_dict(self, values) -> None:
        self.clock.restore(values)
        for key, current in self.device_sums().items():
            setattr(self, key, DeviceArray.from_numpy(np.asarray(values[key], dtype=current.dtype).reshape(-1)))

    def results(self) -> list:
        raise NotImplementedError


@register_accumulator("momentum-distribution")
class MomentumDistribution(Accumulator):
    """Per-mode particle number: the running FP64 sum of |FFT psi|^2, normalised by cell /
    (points x samples), shifted to centre k = 0, less half a quantum per mode with Wigner
    noise; and its g1 (results/accumulators.py:94-127)."""
    name = "momentum-distribution"

    def __init__(self, controller, **params):
        super().__init__(controller, **params)
        from .solvers.split_step_hip import FFTPlan
        g = controller.grid
        self.plan = FFTPlan(g.nx, g.ny, g.dx, g.dy, controller.state.precision)
        self.planes = controller.state.batch_size * controller.state.components
        self.sum_k = DeviceArray.zeros((self.frame.points,), np.float64)
        self._work = DeviceArray.empty((self.frame.points,), controller.state.field.dtype)

    def device_sums(self):
        return {"sum_k": self.sum_k}

    def update(self, state, t) -> None:
        s = self.stream.handle
        N.call("pgpe_fft2", self.plan.handle, state.field.ptr, self._work.ptr, self.planes, 0, s)
        N.call("pgpe_accumulate_power", self._work.ptr, self.sum_k.ptr, self.frame.points, self.frame.precision, s)
        self.stream.synchronize()
        self.clock.tick(t)

    def results(self):
        f = self.frame
        per_mode = self.sum_k.get().reshape(f.shape) * (f.cell / (f.cells * self.samples))
        value = np.fft.fftshift(per_mode, axes=(-2, -1)) - (f.strength / 2 if f.wigner else 0.0)
        k_dims = tuple({"x": "kx", "y": "ky"}.get(d, d) for d in f.dims)
        return [AccumulatedResult("momentum_distribution", value, k_dims, "particles/mode",
                                  "Momentum distribution", f.wigner, True),
                AccumulatedResult("g1", g1(value, ndim=self.ndim), f.dims, "",
                                  "First-order spatial correlation", f.wigner)]


@register_accumulator("g2")
class G2Accumulator(Accumulator):
    """Running FP64 sums of rho and rho^2; the results are their means (normally ordered with
    Wigner noise: <rho^2> - 2 d <rho> + d^2 / 2 and <rho> - d / 2, d = strength / cell) and
    their ratio g2 (results/accumulators.py:130-170)."""
    name = "g2"

    def __init__(self, controller, **params):
        super().__init__(controller, **params)
        self.sum_density = DeviceArray.zeros((self.frame.points,), np.float64)
        self.sum_density_squared = DeviceArray.zeros((self.frame.points,), np.float64)

    def device_sums(self):
        return {"sum_density": self.sum_density, "sum_density_squared": self.sum_density_squared}

    def update(self, state, t) -> None:
        N.call("pgpe_accumulate_moments", state.field.ptr, self.sum_density.ptr, self.sum_density_squared.ptr,
               self.frame.points, self.frame.precision, self.stream.handle)
        self.stream.synchronize()
        self.clock.tick(t)

    def results(self):
        f = self.frame
        rho = self.sum_density.get().reshape(f.shape) / self.samples
        rho2 = self.sum_density_squared.get().reshape(f.shape) / self.samples
        if f.wigner:
            d = f.delta
            rho, rho2 = rho - d / 2, rho2 - 2 * d * rho + d * d / 2
        units = f"µm^-{self.ndim}"
        return [AccumulatedResult("density_mean", rho, f.dims, units, "Mean density", f.wigner, True),
                AccumulatedResult("density_squared_mean", rho2, f.dims, f"µm^-{2 * self.ndim}",
                                  "Mean squared density", f.wigner, True),
                AccumulatedResult("g2", g2(rho, rho2), f.dims, "", "Second-order correlation", f.wigner)]


@dataclass
class WelchRing:
    """Bookkeeping of overlapping windowed segments over a ring of ``segment`` line slots: a
    new segment is due once the ring has filled and, after the first, every half segment of
    new lines (results/accumulators.py:173-260 schedule).  Checkpoints store ``position``,
    ``since_segment`` and ``segments``."""
    segment: int
    position: int = 0         # slot the next line goes to
    since_segment: int = 0    # lines written since the last segment
    segments: int = 0         # segments transformed so far

    def push(self, samples: int) -> bool:
        """Account for the line just written (``samples`` lines in all); True if a segment is due."""
        self.position = (self.position + 1) % self.segment
        self.since_segment += 1
        return samples >= self.segment and (self.segments == 0 or self.since_segment >= self.segment // 2)

    def closed(self) -> None:
        self.segments += 1
        self.since_segment = 0

    def as_arrays(self) -> dict:
        return {k: np.asarray([getattr(self, k)]) for k in ("position", "since_segment", "segments")}

    def restore(self, values) -> None:
        for k in ("position", "since_segment", "segments"):
            setattr(self, k, int(np.asarray(values[k]).reshape(-1)[0]))


class Spectrum(Accumulator):
    """Time-resolved line cut through overlapping Hann windows (results/accumulators.py:173-270).
    Every sample writes one line into a device ring of ``segment`` slots (the k-space cut at
    ky = 0 or the real-space row nearest ``y``); each due segment, the ring in time order times
    the window is transformed along time (complex128) and its power added to the FP64 sums."""
    parameter_keys = frozenset({"segment", "y"})
    axis = ""

    def __init__(self, controller, *, segment=1024, y=0.0, **params):
        super().__init__(controller, **params)
        from .solvers.split_step_hip import FFTPlan
        if not _fft_length_ok(segment):
            raise ValueError(f"{self.name}: segment must be 2^a (<= 4096) or 3 * 2^a (<= 3072), got {segment}")
        state, g = controller.state, controller.grid
        self.welch = WelchRing(int(segment))
        self.segment = self.welch.segment
        self.y = y
        self.row = int(np.argmin(np.abs(g.Y[:, 0] - y))) if self.axis == "x" else None
        self.planes = state.batch_size * state.components
        self.nx = g.nx
        self.width = self.planes * self.nx
        self.ring = DeviceArray.zeros((self.segment * self.width,), state.field.dtype)
        self.psd_sum = DeviceArray.zeros((self.planes * self.segment * self.nx,), np.float64)
        self.window = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(self.segment, dtype=np.float64) / self.segment)
        self._window_dev = DeviceArray.from_numpy(self.window)
        self._block = DeviceArray.empty((self.segment * self.width,), np.complex128)
        self._time_plan = FFTPlan(2, self.segment, 1.0, 1.0, "double")
        if self.axis == "k":
            self._plan = FFTPlan(g.nx, g.ny, g.dx, g.dy, state.precision)
            self._work = DeviceArray.empty((self.frame.points,), state.field.dtype)

    def device_sums(self):
        return {"psd_sum": self.psd_sum, "ring": self.ring}

    def _write_line(self, state, slot, s) -> None:
        f, ny = self.frame, self.controller.grid.ny
        if self.axis == "k":
            N.call("pgpe_fft2", self._plan.handle, state.field.ptr, self._work.ptr, self.planes, 0, s)
            N.call("pgpe_spectrum_line", self._work.ptr, slot.ptr, self.planes, self.nx, ny, 0, 1,
                   float((f.cell / f.cells) ** 0.5), f.precision, s)
        else:
            N.call("pgpe_spectrum_line", state.field.ptr, slot.ptr, self.planes, self.nx, ny, self.row, 0, 1.0,
                   f.precision, s)

    def update(self, state, t) -> None:
        s = self.stream.handle
        self._write_line(state, self.ring.view_range(self.welch.position * self.width, (self.width,)), s)
        self.clock.tick(t)
        if self.welch.push(self.samples):
            N.call("pgpe_spectrum_block", self.ring.ptr, self._block.ptr, self._window_dev.ptr, self.segment,
                   self.width, self.welch.position, self.frame.precision, s)
            N.call("pgpe_fft_columns", self._time_plan.handle, self._block.ptr, self._block.ptr, self.width, 1, s)
            N.call("pgpe_spectrum_power", self._block.ptr, self.psd_sum.ptr, self.segment, self.width, self.nx, s)
            self.welch.closed()
        self.stream.synchronize()

    def state_dict(self):
        return {**super().state_dict(), **self.welch.as_arrays()}

    def load_state_dict(self, values):
        super().load_state_dict(values)
        self.welch.restore(values)

    def results(self):
        import warnings
        if not self.welch.segments:
            warnings.warn(f"{self.name} has no completed segments; no spectrum emitted", UserWarning, stacklevel=2)
            return []
        cfg = self.controller.cfg
        batched = self.controller.state.batched
        spacing = self.every or cfg.result.save_every or cfg.solver.dt
        # energy axis, highest first: E = -2 pi hbar f, plus each member's drive energy
        freq = np.fft.fftfreq(self.segment, d=spacing)
        order = np.argsort(-freq)
        base = -2 * np.pi * cfg.physics.hbar * freq[order]
        shifts = np.asarray([m.drive.energy if m.drive.enabled else 0.0
                             for m in (self.controller._member_cfgs or [cfg])])
        energy = base[None, :] + shifts[:, None] if batched else base + shifts[0]
        d_energy = 2 * np.pi * cfg.physics.hbar / (self.segment * spacing)
        lead = self.frame.shape[:-2]
        psd = self.psd_sum.get().reshape(*lead, self.segment, self.nx)
        norm = self.welch.segments * self.segment * np.sum(self.window ** 2) * d_energy
        value = np.take(psd, order, axis=-2) / norm
        prefix = self.frame.dims[:-2]
        name = f"spectrum_{self.axis}"
        units = "particles/mode/meV" if self.axis == "k" else f"µm^-{self.ndim} meV^-1"
        return [AccumulatedResult(name, value, (*prefix, "E", "kx" if self.axis == "k" else "x"), units,
                                  f"Spectrum {self.axis}", False, True),
                AccumulatedResult(f"{name}_energy", energy, ("member",) * batched + ("E",), "meV", "Energy")]


@register_accumulator("spectrum-k")
class SpectrumK(Spectrum):
    """Momentum cut at ky = 0 (accumulators.py:263-268)."""
    name = "spectrum-k"
    axis = "k"
    parameter_keys = frozenset({"segment"})


@register_accumulator("spectrum-x")
class SpectrumX(Spectrum):
    """Real-space cut at the row nearest y (accumulators.py:271-275)."""
    name = "spectrum-x"
    axis = "x"
