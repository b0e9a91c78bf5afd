"""Result frames: device snapshots, one background writer, NPY / JSON storage (format 3), reader.

Follows results/visitors/storage_visitor.py:34-250 (snapshot, bounded queue, one writer
thread), results/storage/base_storage.py:29-174 (frame buffers, dataset metadata, batching
by frame count and by 256 MiB), results/storage/npy_storage.py:15-136 and json_storage.py
(the two directory layouts of docs/reference/results-format.md) and results/reader.py (the
public reader).  The controller's default result nodes are those of
simulation_controller.py:477-576: density and norm; for spinors density_plus / _minus,
stokes and degree_of_polarization; the reservoir's fields and their maxima when the
reservoir exposes results; psi with ``save_psi``; the pump and the per-laser maxima when
the laser exposes results.

On the device a frame is one ``pgpe_frame_fields`` launch (density, Stokes, degree of
polarisation) plus device copies of psi and the reservoir fields into a slot of a fixed pool,
issued on the recorder's own stream after the solver's advance; the solver's stream waits for
the slot's event before it runs on.  The writer thread copies a slot to the host on a third
stream and feeds every storage backend; the pool bounds the pending frames (at most 64 and
1 GiB of device memory, results/visitors/storage_visitor.py:76-86).

The HDF5 layout (``results.h5``) needs h5py, which is not installed: asking for it raises
``SolverCompatibilityError``.  A resumed run writes a new run directory instead of trimming
and appending to the old one.
"""
from __future__ import annotations

import json
import queue
import threading
import time
from dataclasses import asdict, dataclass, is_dataclass
from pathlib import Path

import numpy as np

from . import _native as N
from . import __version__
from .device import DeviceArray, Event, Stream
from .solver_api import Registry, SolverCompatibilityError

FORMAT_VERSION = 3
MAX_BATCH_BYTES = 256 * 1024 ** 2
_POOL_BYTES = 1024 ** 3
_POOL_FRAMES = 64


@dataclass(frozen=True)
class StorageContext:
    """Member count (batched output), member indices and seeds, resume flag and time
    (results/storage/base_storage.py:14-27)."""
    members: int | None = None
    member_indices: tuple[int, ...] | None = None
    member_seeds: tuple[int, ...] | None = None
    resume: bool = False
    resume_time: float | None = None


@dataclass
class ResultNode:
    """What a frame holds for one saved quantity (results/result_node.py)."""
    name: str
    label: str | None = None
    units: str | None = None
    is_field: bool = True
    static: bool = False
    save: bool = True


def create_run_directory(root) -> Path:
    """A new timestamped directory under ``root`` (storage_visitor.py:19-31)."""
    root = Path(root)
    stamp = time.strftime("%Y%m%d-%H%M%S")
    for suffix in range(10000):
        path = root / (stamp if suffix == 0 else f"{stamp}-{suffix}")
        try:
            path.mkdir(parents=True, exist_ok=False)
            return path
        except FileExistsError:
            continue
    raise RuntimeError(f"could not create a run directory under {root}")


def _jsonable(value):
    if isinstance(value, (np.floating, np.integer)):
        return value.item()
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, tuple):
        return list(value)
    return value


def _config_dict(config):
    return json.loads(json.dumps(asdict(config) if is_dataclass(config) else (config or {}), default=_jsonable))


# ------------------------------------------------------------------------------- storage
class BaseStorage:
    """Frame buffers shared by the backends; a batch is written every ``batch_size`` frames
    or once it holds 256 MiB."""

    suffix = ""

    def __init__(self, output_dir: Path, batch_size: int, config, x, y, context: StorageContext):
        self.output_dir = Path(output_dir)
        self.output_dir.mkdir(parents=True, exist_ok=True)
        self.batch_size = int(batch_size)
        self.ndim = config.grid.ndim if config is not None else (2 if y is not None else 1)
        self.components = 2 if config is not None and config.physics.spin == "spinor" else 1
        self.reservoir_components = ((config.reservoir.components or self.components) if config is not None else 1)
        self.members = context.members
        self.context = context
        self.batch_number = 0
        self.batch_count = 0
        self.batch_bytes = 0
        self.total_saved = 0
        self.time_buffer: list[float] = []
        self.field_buffers: dict[str, list[np.ndarray]] = {}
        self.scalar_buffers: dict[str, list] = {}
        self.scalar_group_buffers: dict[str, dict[str, list]] = {}
        self.static_fields: dict[str, np.ndarray] = {}
        self.node_metadata: dict[str, dict] = {}
        self.metadata = {"format_version": FORMAT_VERSION, "polarism_version": __version__, "ndim": self.ndim,
                         "datasets": {"time": {"label": "time", "units": "ps", "dims": ["time"]}},
                         "config": _config_dict(config)}
        if self.members is not None:
            self.metadata["members"] = self.members
            self.metadata["member_seeds"] = (list(context.member_seeds) if context.member_seeds is not None
                                             else None)
            self.metadata["datasets"]["member_index"] = {"dims": ["member"], "units": ""}

    # public dims and units of a saved array (base_storage.py:60-83)
    def dataset_metadata(self, name: str, *, static: bool = False, field: bool = False) -> dict:
        axes = ["y", "x"] if self.ndim == 2 else ["x"]
        if name == "momentum_density" or (self.components == 2 and name in {"momentum_density_plus",
                                                                             "momentum_density_minus"}):
            axes = ["k" + a for a in axes]
        dims = [] if static else ["time"]
        if self.members is not None:
            dims.append("member")
        if field:
            if name == "stokes":
                dims.append("stokes")
            elif self.components == 2 and (
                    name in {"psi", "density", "momentum_density", "current_x", "current_y", "vortex_charge", "drive"}
                    or (name in {"nR", "nA", "nI", "N_R", "N_I"} and self.reservoir_components == 2)):
                dims.append("component")
            dims.extend(axes)
        units = self.node_metadata.get(name, {}).get("units", "")
        if (name == "density" or (self.components == 2 and name in {"density_plus", "density_minus", "stokes"})
                or name.startswith(("N_R", "nR", "nA", "nI"))):
            units = f"µm^-{self.ndim}"
        elif name == "psi":
            units = f"µm^-{self.ndim / 2:g}"
        elif name == "potential":
            units = "meV"
        return {**self.node_metadata.get(name, {"label": name}), "dims": dims, "units": units}

    def add_to_batch(self, t: float, nodes: list[ResultNode], cached: dict, scalar_groups: dict) -> None:
        self.time_buffer.append(float(t))
        for node in nodes:
            if not node.save or node.name not in cached:
                continue
            self.node_metadata[node.name] = {"label": node.label or node.name,
                                             **({"units": node.units} if node.units is not None else {})}
            value = cached[node.name]
            if node.is_field:
                if node.static:
                    if node.name not in self.static_fields:
                        static = value
                        if (self.members is not None and value.ndim == self.ndim + 1
                                and all(np.array_equal(value[0], v) for v in value[1:])):
                            static = value[0]
                        self.static_fields[node.name] = static
                        self.write_static(node.name, static)
                    continue
                self.field_buffers.setdefault(node.name, []).append(value)
                self.batch_bytes += value.nbytes
            else:
                if self.members is not None and np.shape(value) != (self.members,):
                    raise ValueError(f"batched scalar {node.name!r} must have one value per member")
                self.scalar_buffers.setdefault(node.name, []).append(
                    np.asarray(value) if self.members is not None else float(value))
        for group, values in scalar_groups.items():
            buf = self.scalar_group_buffers.setdefault(group, {})
            for label, value in values.items():
                if self.members is not None and np.shape(value) != (self.members,):
                    raise ValueError(f"batched scalar group {group}/{label} must have one value per member")
                buf.setdefault(label, []).append(np.asarray(value) if self.members is not None else float(value))
        self.batch_count += 1
        if self.batch_count >= self.batch_size or self.batch_bytes >= MAX_BATCH_BYTES:
            self.dump_batch()

    def _record_datasets(self) -> None:
        for name in self.field_buffers:
            self.metadata["datasets"][f"fields/{name}"] = self.dataset_metadata(name, field=True)
        for name in self.scalar_buffers:
            self.metadata["datasets"][f"scalars/{name}"] = self.dataset_metadata(name)
        for group, values in self.scalar_group_buffers.items():
            for label in values:
                self.metadata["datasets"][f"scalar_groups/{group}/{label}"] = self.dataset_metadata(label)

    def _static_metadata(self, name: str, value: np.ndarray) -> None:
        meta = self.dataset_metadata(name, static=True, field=True)
        if self.members is not None and value.ndim == self.ndim:
            meta["dims"].remove("member")
        self.metadata["datasets"][f"static/{name}"] = meta

    def _write_metadata(self) -> None:
        (self.output_dir / "metadata.json").write_text(json.dumps(self.metadata, sort_keys=True, default=_jsonable))

    def _clear(self) -> None:
        self.total_saved += len(self.time_buffer)
        self.time_buffer = []
        self.field_buffers = {}
        self.scalar_buffers = {}
        self.scalar_group_buffers = {}
        self.batch_count = 0
        self.batch_bytes = 0

    def flush(self) -> None:
        if self.batch_count > 0:
            self.dump_batch()

    def finalize(self) -> None:
        self.flush()
        self._write_metadata()

    def set_stop_times(self, stop_times) -> None:
        self.metadata["stop_time"] = np.asarray(stop_times, dtype=float).tolist()

    def dump_batch(self) -> None:
        raise NotImplementedError

    def write_static(self, name: str, value: np.ndarray) -> None:
        raise NotImplementedError

    def write_accumulated(self, results: dict, metadata: dict) -> None:
        raise NotImplementedError


available_storages: Registry[type] = Registry("result storage")


def register_storage(name: str):
    return available_storages.register(name)


@register_storage("npy")
class NPYStorage(BaseStorage):
    """``metadata.json``, ``batch_NNNNNN.npz`` (time, field_<name>, scalar_<name>,
    group_<group>_<label>), ``static_<name>.npy``, ``grid_x.npy`` / ``grid_y.npy``,
    ``member_index.npy``, ``accumulated_<name>.npy`` (npy_storage.py:15-136)."""

    def __init__(self, output_dir, batch_size, config, x, y, context):
        super().__init__(output_dir, batch_size, config, x, y, context)
        self.compress = config is not None and config.result.compression != "none"
        if self.members is not None:
            np.save(self.output_dir / "member_index.npy", np.asarray(context.member_indices))
        if x is not None:
            np.save(self.output_dir / "grid_x.npy", np.asarray(x))
            self.metadata["datasets"]["grid/x"] = {"units": "µm", "dims": ["x"]}
        if y is not None:
            np.save(self.output_dir / "grid_y.npy", np.asarray(y))
            self.metadata["datasets"]["grid/y"] = {"units": "µm", "dims": ["y"]}

    def dump_batch(self) -> None:
        if not self.time_buffer:
            return
        arrays = {"time": np.asarray(self.time_buffer, dtype=np.float64)}
        for name, frames in self.field_buffers.items():
            arrays[f"field_{name}"] = np.stack(frames, axis=0)
        for name, values in self.scalar_buffers.items():
            arrays[f"scalar_{name}"] = np.asarray(values)
        for group, values in self.scalar_group_buffers.items():
            for label, v in values.items():
                arrays[f"group_{group}_{label}"] = np.asarray(v)
        path = self.output_dir / f"batch_{self.batch_number:06d}.npz"
        (np.savez_compressed if self.compress else np.savez)(path, **arrays)
        self.batch_number += 1
        self._record_datasets()
        if self.batch_number == 1:
            self._write_metadata()
        self._clear()

    def write_static(self, name, value):
        np.save(self.output_dir / f"static_{name}.npy", value)
        self._static_metadata(name, value)

    def write_accumulated(self, results, metadata):
        self.metadata["accumulated"] = {}
        for name, r in results.items():
            np.save(self.output_dir / f"accumulated_{name}.npy", np.asarray(r.value))
            self.metadata["accumulated"][name] = {**metadata[name], "dims": list(r.dims), "units": r.units,
                                                  "label": r.label, "wigner_corrected": r.wigner_corrected,
                                                  "linear": r.linear}


@register_storage("json")
class JSONStorage(BaseStorage):
    """``metadata.json`` and ``batch_NNNNNN.json`` with time / fields / scalars /
    scalar_groups (static fields in the first batch); complex arrays as
    ``{"real": ..., "imag": ...}`` (json_storage.py)."""

    def __init__(self, output_dir, batch_size, config, x, y, context):
        super().__init__(output_dir, batch_size, config, x, y, context)
        if self.members is not None:
            (self.output_dir / "member_index.json").write_text(json.dumps(list(context.member_indices)))
            if x is not None:
                (self.output_dir / "grid_x.json").write_text(json.dumps(np.asarray(x).tolist()))
                self.metadata["datasets"]["grid/x"] = {"units": "µm", "dims": ["x"]}
            if y is not None:
                (self.output_dir / "grid_y.json").write_text(json.dumps(np.asarray(y).tolist()))
                self.metadata["datasets"]["grid/y"] = {"units": "µm", "dims": ["y"]}

    @staticmethod
    def _field(value: np.ndarray):
        if np.iscomplexobj(value):
            return {"real": value.real.tolist(), "imag": value.imag.tolist()}
        return value.tolist()

    def dump_batch(self) -> None:
        if not self.time_buffer:
            return
        batch = {"time": list(self.time_buffer), "fields": {}, "scalars": {}, "scalar_groups": {}}
        if self.batch_number == 0 and self.static_fields:
            batch["static"] = {name: self._field(v) for name, v in self.static_fields.items()}
        for name, frames in self.field_buffers.items():
            batch["fields"][name] = [self._field(f) for f in frames]
        for name, values in self.scalar_buffers.items():
            batch["scalars"][name] = np.asarray(values).tolist()
        for group, values in self.scalar_group_buffers.items():
            batch["scalar_groups"][group] = {label: np.asarray(v).tolist() for label, v in values.items()}
        with open(self.output_dir / f"batch_{self.batch_number:06d}.json", "w") as f:
            json.dump(batch, f, indent=2)
        self.batch_number += 1
        self._record_datasets()
        if self.batch_number == 1:
            self._write_metadata()
        self._clear()

    def write_static(self, name, value):
        self._static_metadata(name, value)

    def write_accumulated(self, results, metadata):
        self.metadata["accumulated"] = {
            name: {**metadata[name], "dims": list(r.dims), "units": r.units, "label": r.label,
                   "wigner_corrected": r.wigner_corrected, "linear": r.linear, "value": self._field(np.asarray(r.value))}
            for name, r in results.items()}


def storage_names(result_cfg) -> list[str]:
    """Formats asked for by ``result.formats`` and the save_* flags, in order, once each."""
    names = dict.fromkeys((*result_cfg.formats,
                           *(n for on, n in ((result_cfg.save_hdf5, "hdf5"), (result_cfg.save_npy, "npy"),
                                             (result_cfg.save_json, "json")) if on)))
    for name in names:
        if name == "hdf5":
            raise SolverCompatibilityError("result format 'hdf5' needs h5py, which is not installed here; "
                                           "use result.formats = ('npy',) or ('json',)")
        available_storages.lookup(name)
    return list(names)


# ------------------------------------------------------------------------------- frames
class FrameRecorder:
    """Snapshot the controller's result nodes on the device and write them on one thread."""

    def __init__(self, ctl, output_dir=None, context: StorageContext | None = None):
        cfg = ctl.cfg
        self.ctl = ctl
        self.cfg = cfg
        self.context = context or StorageContext()
        self.output_dir = Path(output_dir) if output_dir is not None else create_run_directory(
            cfg.result.output_directory)
        self.output_dir.mkdir(parents=True, exist_ok=True)
        grid = ctl.grid
        x = np.asarray(grid.X[0, :])
        y = np.asarray(grid.Y[:, 0])
        self.storages = [available_storages.lookup(name)(self.output_dir, cfg.result.batch_size, cfg, x, y,
                                                          self.context)
                         for name in storage_names(cfg.result)]
        self.nodes = self._build_nodes()
        self._static_seen: set[str] = set()
        st = ctl.state
        self.batched = st.batched
        self.members = st.batch_size
        self.components = st.components
        self.plane = int(np.prod(grid.shape))
        self.real = np.float32 if cfg.solver.precision == "single" else np.float64
        self.prec = 0 if self.real is np.float32 else 1
        self.res_names = tuple(getattr(ctl.reservoir, "state_names", ()))
        self.res_comps = getattr(ctl.reservoir, "components", 1)
        self._stream = Stream()
        self._free: queue.Queue | None = None
        self._work: queue.Queue | None = None
        self._thread = None
        self._error: BaseException | None = None
        self._closed = False

    # nodes of simulation_controller.py:477-576
    def _build_nodes(self) -> list[ResultNode]:
        ctl, cfg = self.ctl, self.cfg
        nodes = [ResultNode("density", "|ψ|²"), ResultNode("norm", "N(t)", is_field=False)]
        if ctl.state.components == 2:
            nodes += [ResultNode("density_plus"), ResultNode("density_minus"), ResultNode("stokes"),
                      ResultNode("degree_of_polarization")]
        if cfg.reservoir.expose_results:
            public = {"single": ("N_R",), "double": ("nA", "nI"), "quadratic-double": ("nR", "nI")}.get(
                ctl.reservoir.kind, ())
            nodes += [ResultNode(n) for n in public]
            nodes += [ResultNode(n + "_max", is_field=False) for n in public]
        if cfg.result.save_psi:
            nodes.append(ResultNode("psi"))
        if cfg.laser.expose_results:
            lasers = self._lasers()
            static = all(getattr(l, "time_independent", True) and l.delay <= 0 for l in lasers)
            nodes.append(ResultNode("pump", "Pump", static=static))
        return nodes

    def _lasers(self):
        pump = self.ctl.pump
        if hasattr(pump, "member_lasers"):
            return [ls[0] for ls in pump.member_lasers]
        return list(pump.lasers)

    def _needed(self) -> list[ResultNode]:
        return [n for n in self.nodes if n.save and not (n.static and n.name in self._static_seen)]

    def _start(self, slot_bytes: int) -> None:
        count = max(1, min(_POOL_FRAMES, _POOL_BYTES // max(slot_bytes, 1)))
        self._free = queue.Queue()
        self._work = queue.Queue()
        for _ in range(count):
            self._free.put(self._new_slot())
        self._thread = threading.Thread(target=self._write_loop, name="pgpe-results", daemon=True)
        self._thread.start()

    def _new_slot(self) -> dict:
        B, C, shp = self.members, self.components, self.ctl.grid.shape
        slot = {"density": DeviceArray.empty((B, C, *shp), self.real), "event": Event()}
        if C == 2:
            slot["stokes"] = DeviceArray.empty((B, 4, *shp), self.real)
            slot["dop"] = DeviceArray.empty((B, *shp), self.real)
        if self.cfg.result.save_psi:
            slot["psi"] = DeviceArray.empty(self.ctl.state.field.shape, self.ctl.state.field.dtype)
        if self.cfg.reservoir.expose_results:
            for name, arr in zip(self.res_names, self.ctl.reservoir.get_state()):
                slot["res_" + name] = DeviceArray.empty(arr.shape, arr.dtype)
        return slot

    def _slot_bytes(self) -> int:
        B, C = self.members, self.components
        n = B * C * self.plane * np.dtype(self.real).itemsize
        if C == 2:
            n += 5 * B * self.plane * np.dtype(self.real).itemsize
        if self.cfg.result.save_psi:
            n += self.ctl.state.field.nbytes
        if self.cfg.reservoir.expose_results:
            n += sum(a.nbytes for a in self.ctl.reservoir.get_state())
        return n

    def _raise_error(self) -> None:
        if self._error is not None:
            raise RuntimeError("result writer failed") from self._error

    def _take_slot(self) -> dict:
        while True:
            self._raise_error()
            try:
                return self._free.get(timeout=0.05)
            except queue.Empty:
                continue

    def record(self, t: float) -> None:
        """Snapshot the current state (after an advance) into a device slot and queue it."""
        self._raise_error()
        if self._closed:
            raise RuntimeError("result recorder is closed")
        if self._free is None:
            self._start(self._slot_bytes())
        nodes = self._needed()
        slot = self._take_slot()
        s = self._stream.handle
        field = self.ctl.state.field
        C = self.components
        N.call("pgpe_frame_fields", field.ptr, self.prec, C, self.plane, self.members, slot["density"].ptr,
               slot["stokes"].ptr if C == 2 else None, slot["dop"].ptr if C == 2 else None, s)
        if "psi" in slot:
            slot["psi"].copy_from(field, s)
        if self.cfg.reservoir.expose_results:
            for name, arr in zip(self.res_names, self.ctl.reservoir.get_state()):
                slot["res_" + name].copy_from(arr, s)
        slot["event"].record(s)
        # the solver's next launches wait for the snapshot
        solver_stream = getattr(self.ctl.solver, "stream", None)
        if solver_stream is not None:
            N.call("pgpe_stream_wait_event", solver_stream.handle, slot["event"].handle)
        else:
            self._stream.synchronize()
        host_extra, groups = self._host_values(t, nodes)
        for n in nodes:
            if n.static:
                self._static_seen.add(n.name)
        self._work.put((slot, float(t), nodes, host_extra, groups))

    def _host_values(self, t, nodes):
        """Pump field and per-laser maxima (host-side fields, simulation_controller.py:1185-1238)."""
        extra, groups = {}, {}
        names = {n.name for n in nodes}
        if "pump" in names:
            P = np.asarray(self.ctl.pump(t))
            if self.batched and P.ndim == 4:
                P = P[:, 0]
            extra["pump"] = np.array(P, copy=True)
        if self.cfg.laser.expose_results:
            if self.batched:
                lasers = self._lasers()
                groups["P_lasers"] = {"L0": np.asarray([np.max(l.envelope * l.power_coefficient(t))
                                                        for l in lasers])}
            else:
                groups["P_lasers"] = {f"L{i}": float(np.max(l.envelope * l.power_coefficient(t)))
                                      for i, l in enumerate(self._lasers())}
        return extra, groups

    def _public(self, a: np.ndarray) -> np.ndarray:
        """(B, C, *S) -> the node's public shape: no member axis unbatched, no component axis
        for one component."""
        if a.shape[1] == 1:
            a = a[:, 0]
        return a if self.batched else a[0]

    def _frame(self, slot, nodes, host_extra, copy_stream):
        names = {n.name for n in nodes}
        get = lambda key: slot[key].get(stream=copy_stream)  # noqa: E731
        cached = {}
        dens = get("density")                                   # (B, C, *S)
        cell = float(np.prod(self.ctl.grid.spacing))
        if "density" in names:
            cached["density"] = self._public(dens)
        if "norm" in names:
            norm = dens.reshape(self.members, -1).sum(axis=1, dtype=np.float64) * cell
            cached["norm"] = norm if self.batched else float(norm[0])
        if self.components == 2:
            if "density_plus" in names:
                cached["density_plus"] = dens[:, 0] if self.batched else dens[0, 0]
            if "density_minus" in names:
                cached["density_minus"] = dens[:, 1] if self.batched else dens[0, 1]
            if "stokes" in names:
                st = get("stokes")
                cached["stokes"] = st if self.batched else st[0]
            if "degree_of_polarization" in names:
                dp = get("dop")
                cached["degree_of_polarization"] = dp if self.batched else dp[0]
        if "psi" in names:
            cached["psi"] = self._public(get("psi"))
        if self.cfg.reservoir.expose_results:
            public = {"single": ("N_R",), "double": ("nA", "nI"), "quadratic-double": ("nR", "nI")}.get(
                self.ctl.reservoir.kind, ())
            for pub, name in zip(public, self.res_names):
                if pub not in names and pub + "_max" not in names:
                    continue
                r = self._public(get("res_" + name))
                if pub in names:
                    cached[pub] = r
                if pub + "_max" in names:
                    if self.batched:
                        both = self.components == 2 and self.res_comps == 2
                        axes = tuple(range(-self.ctl.grid.ndim - (1 if both else 0), 0))
                        cached[pub + "_max"] = r.max(axis=axes)
                    else:
                        cached[pub + "_max"] = float(r.max())
        cached.update(host_extra)
        return cached

    def _write_loop(self) -> None:
        N.call("pgpe_device_init", int(self.cfg.compute_engine.gpu_device))  # the current device is per thread
        copy_stream = Stream()
        while True:
            item = self._work.get()
            try:
                if item is None:
                    return
                slot, t, nodes, host_extra, groups = item
                if self._error is None:
                    N.call("pgpe_stream_wait_event", copy_stream.handle, slot["event"].handle)
                    cached = self._frame(slot, nodes, host_extra, copy_stream)
                    self._free.put(slot)
                    for storage in self.storages:
                        storage.add_to_batch(t, nodes, cached=cached, scalar_groups=groups)
                else:
                    self._free.put(slot)
            except BaseException as exc:  # surfaced on the caller's next call
                self._error = exc
                if item is not None:
                    self._free.put(item[0])
            finally:
                self._work.task_done()

    def flush(self) -> None:
        """Drain the queued frames and write every backend's current batch."""
        if self._work is not None and not self._closed:
            self._work.join()
        self._raise_error()
        for storage in self.storages:
            storage.flush()

    def set_stop_times(self, stop_times) -> None:
        self.flush()
        for storage in self.storages:
            storage.set_stop_times(stop_times)

    def set_accumulated(self, results: dict, metadata: dict) -> None:
        self.flush()
        for storage in self.storages:
            storage.write_accumulated(results, metadata)

    def finalize(self) -> None:
        if self._closed:
            self._raise_error()
            return
        if self._work is not None:
            self._work.put(None)
            self._thread.join()
        self._closed = True
        try:
            for storage in self.storages:
                storage.finalize()
        finally:
            self._raise_error()


# ------------------------------------------------------------------------------- reader
class Results:
    """One run's saved results (results/reader.py): ``time``, ``fields``, ``scalars``,
    ``scalar_groups``, ``static``, ``accumulated``, ``accumulated_meta``, ``config``, ``x``,
    ``y``, ``attrs`` (the metadata), ``kx`` / ``ky`` (fftshifted angular wave numbers, rad/µm)."""

    def __init__(self, path: Path, attrs: dict, time, fields, scalars, groups, static, accumulated, x, y):
        self.path = path
        self.attrs = attrs
        self.config = attrs.get("config", {})
        self.time = time
        self.fields = fields
        self.scalars = scalars
        self.scalar_groups = groups
        self.static = static
        self.accumulated = accumulated
        self.accumulated_meta = {k: {kk: vv for kk, vv in v.items() if kk != "value"}
                                 for k, v in attrs.get("accumulated", {}).items()}
        self.x = x
        self.y = y

    @staticmethod
    def _k(coord):
        if coord is None or len(coord) < 2:
            return None
        d = float(coord[1] - coord[0])
        return np.fft.fftshift(2 * np.pi * np.fft.fftfreq(len(coord), d=d))

    @property
    def kx(self):
        return self._k(self.x)

    @property
    def ky(self):
        return self._k(self.y)

    def close(self) -> None:
        pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _concat(parts: dict[str, list]) -> dict:
    return {k: np.concatenate(v, axis=0) for k, v in parts.items()}


def _load_npy(path: Path) -> Results:
    attrs = json.loads((path / "metadata.json").read_text())
    times, fields, scalars, groups = [], {}, {}, {}
    for batch in sorted(path.glob("batch_*.npz")):
        with np.load(batch, allow_pickle=False) as z:
            times.append(z["time"])
            for key in z.files:
                if key.startswith("field_"):
                    fields.setdefault(key[6:], []).append(z[key])
                elif key.startswith("scalar_"):
                    scalars.setdefault(key[7:], []).append(z[key])
                elif key.startswith("group_"):
                    groups.setdefault(key[6:], []).append(z[key])
    # split group_<group>_<label> by the datasets the metadata lists
    known = {}
    for key in attrs.get("datasets", {}):
        if key.startswith("scalar_groups/"):
            _, group, label = key.split("/", 2)
            known[f"{group}_{label}"] = (group, label)
    sg: dict[str, dict] = {}
    for key, value in _concat(groups).items():
        group, label = known.get(key, key.rpartition("_")[::2])
        sg.setdefault(group, {})[label] = value
    static = {p.name[7:-4]: np.load(p, allow_pickle=False) for p in path.glob("static_*.npy")}
    acc = {p.name[12:-4]: np.load(p, allow_pickle=False) for p in path.glob("accumulated_*.npy")}
    x = np.load(path / "grid_x.npy") if (path / "grid_x.npy").exists() else None
    y = np.load(path / "grid_y.npy") if (path / "grid_y.npy").exists() else None
    time_axis = np.concatenate(times) if times else np.zeros(0)
    return Results(path, attrs, time_axis, _concat(fields), _concat(scalars), sg, static, acc, x, y)


def _json_array(value):
    if isinstance(value, dict) and set(value) == {"real", "imag"}:
        return np.asarray(value["real"]) + 1j * np.asarray(value["imag"])
    if isinstance(value, list) and value and isinstance(value[0], dict):
        return np.stack([_json_array(v) for v in value])
    return np.asarray(value)


def _load_json(path: Path) -> Results:
    attrs = json.loads((path / "metadata.json").read_text())
    times, fields, scalars, groups, static = [], {}, {}, {}, {}
    for batch in sorted(path.glob("batch_*.json")):
        data = json.loads(batch.read_text())
        times.append(np.asarray(data["time"], dtype=np.float64))
        for name, frames in data.get("fields", {}).items():
            fields.setdefault(name, []).append(_json_array(frames))
        for name, values in data.get("scalars", {}).items():
            scalars.setdefault(name, []).append(np.asarray(values))
        for group, labels in data.get("scalar_groups", {}).items():
            for label, values in labels.items():
                groups.setdefault(group, {}).setdefault(label, []).append(np.asarray(values))
        for name, value in data.get("static", {}).items():
            static[name] = _json_array(value)
    acc = {name: _json_array(meta["value"]) for name, meta in attrs.get("accumulated", {}).items()
           if "value" in meta}
    gx, gy = path / "grid_x.json", path / "grid_y.json"
    x = np.asarray(json.loads(gx.read_text())) if gx.exists() else None
    y = np.asarray(json.loads(gy.read_text())) if gy.exists() else None
    sg = {g: _concat(v) for g, v in groups.items()}
    time_axis = np.concatenate(times) if times else np.zeros(0)
    return Results(path, attrs, time_axis, _concat(fields), _concat(scalars), sg, static, acc, x, y)


def load_results(path) -> Results:
    """Open a run directory written by the NPY or JSON backend (NPY preferred, as the reference
    prefers HDF5, then NPY, then JSON)."""
    path = Path(path)
    if path.is_file():
        path = path.parent
    if (path / "results.h5").exists():
        raise SolverCompatibilityError("reading results.h5 needs h5py, which is not installed here")
    if (path / "metadata.json").exists() and any(path.glob("batch_*.npz")):
        return _load_npy(path)
    if (path / "metadata.json").exists() and any(path.glob("batch_*.json")):
        return _load_json(path)
    raise FileNotFoundError(f"no NPY or JSON results in {path}")
