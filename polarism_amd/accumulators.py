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


class Accumulator:
    """Scheduled device-side accumulator (accumulators.py:32-91)."""
    name = ""
    parameter_keys: frozenset = frozenset()

    def __init__(self, controller, *, start=0.0, stop=None, every=None, **params):
        if params:
            raise ValueError(f"Unknown {self.name} parameters: {sorted(params)}")
        self.controller = controller
        self.samples = 0
        self.t_first = None
        self.t_last = None
        self.start = start
        self.stop = stop
        self.every = every
        state = controller.state
        grid = controller.grid
        self.ndim = 2
        self.dims = (("member",) if state.batched else ()) + (("component",) if state.components == 2 else ()) + \
            ("y", "x")
        self.shape = state._public_shape
        self.points = int(np.prod(state._canonical_shape))
        cfg = controller.cfg
        self.wigner_corrected = bool(cfg.noise.enabled and cfg.noise.stochastic == "wigner")
        self.cell = grid.dx * grid.dy
        self.cells = grid.nx * grid.ny
        members = controller._member_cfgs or [cfg]
        strengths = [m.noise.stochastic_strength if m.noise.enabled and m.noise.stochastic == "wigner" else 0.0
                     for m in members]
        per = np.asarray(strengths, dtype=np.float64)
        self.delta = (per.reshape((len(per),) + (1,) * (len(self.shape) - 1)) if state.batched
                      else per.reshape((1,) * len(self.shape))) / self.cell
        self.strength = self.delta * self.cell
        self.precision = 0 if state.precision == "single" else 1
        self.stream = Stream()

    def update(self, state, t) -> None:
        raise NotImplementedError

    def _sampled(self, t):
        if self.samples == 0:
            self.t_first = t
        self.t_last = t
        self.samples += 1

    def state_dict(self) -> dict:
        return {"samples": self.samples, "t_first": self.t_first, "t_last": self.t_last}

    def load_state_dict(self, values) -> None:
        self.samples = int(values["samples"])
        self.t_first = values["t_first"]
        self.t_last = values["t_last"]

    def results(self) -> list:
        raise NotImplementedError


@register_accumulator("momentum-distribution")
class MomentumDistribution(Accumulator):
    """Per-mode particle number sum_k |FFT psi|^2 and its g1 (accumulators.py:94-127)."""
    name = "momentum-distribution"

    def __init__(self, controller, **params):
        super().__init__(controller, **params)
        from .solvers.split_step_hip import FFTPlan
        g = controller.grid
        self.plan = FFTPlan(g.nx, g.ny, g.dx, g.dy, controller.state.precision)
        self.planes = controller.state.batch_size * controller.state.components
        self.sum_k = DeviceArray.zeros((self.points,), np.float64)
        self._work = DeviceArray.empty((self.points,), controller.state.field.dtype)

    def update(self, state, t) -> None:
        s = self.stream.handle
        N.call("pgpe_fft2", self.plan.handle, state.field.ptr, self._work.ptr, self.planes, 0, s)
        N.call("pgpe_accumulate_power", self._work.ptr, self.sum_k.ptr, self.points, self.precision, s)
        self.stream.synchronize()
        self._sampled(t)

    def state_dict(self):
        return {**super().state_dict(), "sum_k": self.sum_k.get()}

    def load_state_dict(self, values):
        super().load_state_dict(values)
        self.sum_k = DeviceArray.from_numpy(np.asarray(values["sum_k"], dtype=np.float64).reshape(-1))

    def results(self):
        sum_k = self.sum_k.get().reshape(self.shape)
        value = np.fft.fftshift(sum_k * (self.cell / (self.cells * self.samples)), axes=(-2, -1))
        if self.wigner_corrected:
            value = value - self.strength / 2
        dims = tuple("k" + a if a in ("x", "y") else a for a in self.dims)
        return [AccumulatedResult("momentum_distribution", value, dims, "particles/mode",
                                  "Momentum distribution", self.wigner_corrected, True),
                AccumulatedResult("g1", g1(value, ndim=self.ndim), self.dims, "",
                                  "First-order spatial correlation", self.wigner_corrected)]


@register_accumulator("g2")
class G2Accumulator(Accumulator):
    """Normally ordered density moments and their ratio (accumulators.py:130-170)."""
    name = "g2"

    def __init__(self, controller, **params):
        super().__init__(controller, **params)
        self.sum_density = DeviceArray.zeros((self.points,), np.float64)
        self.sum_density_squared = DeviceArray.zeros((self.points,), np.float64)

    def update(self, state, t) -> None:
        N.call("pgpe_accumulate_moments", state.field.ptr, self.sum_density.ptr, self.sum_density_squared.ptr,
               self.points, self.precision, self.stream.handle)
        self.stream.synchronize()
        self._sampled(t)

    def state_dict(self):
        return {**super().state_dict(), "sum_density": self.sum_density.get(),
                "sum_density_squared": self.sum_density_squared.get()}

    def load_state_dict(self, values):
        super().load_state_dict(values)
        self.sum_density = DeviceArray.from_numpy(np.asarray(values["sum_density"], dtype=np.float64).reshape(-1))
        self.sum_density_squared = DeviceArray.from_numpy(
            np.asarray(values["sum_density_squared"], dtype=np.float64).reshape(-1))

    def results(self):
        mean = self.sum_density.get().reshape(self.shape) / self.samples
        squared = self.sum_density_squared.get().reshape(self.shape) / self.samples
        if self.wigner_corrected:
            squared = squared - 2 * self.delta * mean + self.delta * self.delta / 2
            mean = mean - self.delta / 2
        units = f"µm^-{self.ndim}"
        return [AccumulatedResult("density_mean", mean, self.dims, units, "Mean density", self.wigner_corrected, True),
                AccumulatedResult("density_squared_mean", squared, self.dims, f"µm^-{2 * self.ndim}",
                                  "Mean squared density", self.wigner_corrected, True),
                AccumulatedResult("g2", g2(mean, squared), self.dims, "", "Second-order correlation",
                                  self.wigner_corrected)]


class Spectrum(Accumulator):
    """Time-resolved line cut through overlapping Hann windows (accumulators.py:173-270): every
    sample writes one line into a device ring of ``segment`` slots; once the ring is full and
    half a segment has passed since the last one, the ring in time order times the window is
    transformed along time (complex128) and its power added to the FP64 sums."""
    parameter_keys = frozenset({"segment", "y"})
    axis = ""

    def __init__(self, controller, *, segment=1024, y=0.0, **params):
        super().__init__(controller, **params)
        from .solvers.split_step_hip import FFTPlan
        if not _fft_length_ok(segment):
            raise ValueError(f"{self.name}: segment must be 2^a (<= 4096) or 3 * 2^a (<= 3072), got {segment}")
        state, g = controller.state, controller.grid
        self.segment = int(segment)
        self.y = y
        self.position = 0
        self.since_segment = 0
        self.segments = 0
        self.row = None
        if self.axis == "x":
            self.row = int(np.argmin(np.abs(g.Y[:, 0] - y)))
        self.planes = state.batch_size * state.components
        self.nx, self.ny = g.nx, g.ny
        self.width = self.planes * self.nx
        self.cut_shape = (*self.shape[:-2], self.shape[-1])
        self.ring = DeviceArray.zeros((self.segment * self.width,), state.field.dtype)
        self.psd_sum = DeviceArray.zeros((self.planes * self.segment * self.nx,), np.float64)
        idx = np.arange(self.segment, dtype=np.float64)
        self.window = 0.5 - 0.5 * np.cos(2 * np.pi * idx / self.segment)
        self._window_dev = DeviceArray.from_numpy(self.window)
        self._block = DeviceArray.empty((self.segment * self.width,), np.complex128)
        self._time_plan = FFTPlan(2, self.segment, 1.0, 1.0, "double")
        if self.axis == "k":
            self._plan = FFTPlan(g.nx, g.ny, g.dx, g.dy, state.precision)
            self._work = DeviceArray.empty((self.points,), state.field.dtype)

    def update(self, state, t) -> None:
        s = self.stream.handle
        slot = self.ring.view_range(self.position * self.width, (self.width,))
        if self.axis == "k":
            N.call("pgpe_fft2", self._plan.handle, state.field.ptr, self._work.ptr, self.planes, 0, s)
            N.call("pgpe_spectrum_line", self._work.ptr, slot.ptr, self.planes, self.nx, self.ny, 0, 1,
                   float((self.cell / self.cells) ** 0.5), self.precision, s)
        else:
            N.call("pgpe_spectrum_line", state.field.ptr, slot.ptr, self.planes, self.nx, self.ny, self.row, 0, 1.0,
                   self.precision, s)
        self.position = (self.position + 1) % self.segment
        self.since_segment += 1
        self._sampled(t)
        if self.samples >= self.segment and (self.segments == 0 or self.since_segment >= self.segment // 2):
            N.call("pgpe_spectrum_block", self.ring.ptr, self._block.ptr, self._window_dev.ptr, self.segment,
                   self.width, self.position, self.precision, s)
            N.call("pgpe_fft_columns", self._time_plan.handle, self._block.ptr, self._block.ptr, self.width, 1, s)
            N.call("pgpe_spectrum_power", self._block.ptr, self.psd_sum.ptr, self.segment, self.width, self.nx, s)
            self.segments += 1
            self.since_segment = 0
        self.stream.synchronize()

    def state_dict(self):
        return {**super().state_dict(), "psd_sum": self.psd_sum.get(), "ring": self.ring.get(),
                "position": np.asarray([self.position]), "since_segment": np.asarray([self.since_segment]),
                "segments": np.asarray([self.segments])}

    def load_state_dict(self, values):
        super().load_state_dict(values)
        self.psd_sum = DeviceArray.from_numpy(np.asarray(values["psd_sum"], dtype=np.float64).reshape(-1))
        self.ring = DeviceArray.from_numpy(np.asarray(values["ring"], dtype=self.ring.dtype).reshape(-1))
        self.position = int(values["position"][0])
        self.since_segment = int(values["since_segment"][0])
        self.segments = int(values["segments"][0])

    def results(self):
        import warnings
        if not self.segments:
            warnings.warn(f"{self.name} has no completed segments; no spectrum emitted", UserWarning, stacklevel=2)
            return []
        cfg = self.controller.cfg
        hbar = cfg.physics.hbar
        spacing = self.every or cfg.result.save_every or cfg.solver.dt
        frequencies = np.fft.fftfreq(self.segment, d=spacing)
        order = np.argsort(-frequencies)
        delta_energy = 2 * np.pi * hbar / (self.segment * spacing)
        energy = -2 * np.pi * hbar * frequencies[order]
        batched = self.controller.state.batched
        members = self.controller._member_cfgs or [cfg]
        shifts = [m.drive.energy if m.drive.enabled else 0.0 for m in members]
        energy = energy[None, :] + np.asarray(shifts)[:, None] if batched else energy + shifts[0]
        window_power = np.sum(self.window ** 2)
        psd = self.psd_sum.get().reshape(*self.cut_shape[:-1], self.segment, self.nx)
        value = np.take(psd, order, axis=-2) / (self.segments * self.segment * window_power * delta_energy)
        prefix = (("member",) if batched else ()) + (("component",) if self.controller.state.components == 2 else ())
        name = f"spectrum_{self.axis}"
        units = "particles/mode/meV" if self.axis == "k" else f"µm^-{self.ndim} meV^-1"
        return [AccumulatedResult(name, value, (*prefix, "E", "kx" if self.axis == "k" else "x"), units,
                                  f"Spectrum {self.axis}", False, True),
                AccumulatedResult(f"{name}_energy", energy, (("member",) if batched else ()) + ("E",), "meV",
                                  "Energy")]


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
