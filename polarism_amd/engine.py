"""RK4 launch engine over a packed member table.

Owns the buffers of one batch (uniform or uneven grids) and issues either the
four fused ``pgpe_rk4_stage`` launches of a step or, with ``whole_step``, one
``pgpe_rk4_step`` launch that runs all four stages on halo tiles (same
result, bit for bit on periodic grids; scalar fields; the reservoir then ping-pongs between two
buffers like the field), on 64 x tile_h halo tiles staged in LDS or, with ``strip_h``, on row
strips of 56 output columns held in registers.  Used by the ``rk4-fdm-hip`` solver and by the
packed sweep driver.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .device import DeviceArray, Graph, Stream
from .layout import MemberTable

STRIP_MAX_POINTS = 1 << 28   # per member on row strips: 32-bit byte offsets of complex128 values


def real_dtype(precision):
    return np.float32 if precision == "single" else np.float64


def complex_dtype(precision):
    return np.complex64 if precision == "single" else np.complex128


def spinor_whole_step_applies(table, precision: str) -> bool:
    """Whether the spinor whole-step kernel runs this member table: single precision (double
    needs more registers than its workgroup has), tiles of height 16 and periodic grids only (on
    a closed interval the TE-TM cross term changes sign under the wall's reflection, so the halo
    cannot be recomputed there; the stage kernels mirror one point per stage instead)."""
    return (precision == "single" and table.tile_h == 16
            and all(getattr(s, "boundary", "periodic") == "periodic" for s in table.specs))


class Rk4Engine:
    def __init__(self, table: MemberTable, pump: DeviceArray, precision: str, reservoir: str,
                 stream: Stream | None = None, whole_step: bool = False, strip_h: int | None = None):
        self.table = table
        self.pump = pump
        self.precision = precision
        self.prec_code = 0 if precision == "single" else 1
        self.rdtype = real_dtype(precision)
        self.cdtype = complex_dtype(precision)
        self.reservoir = reservoir
        self.nres = N.RESERVOIR_FIELDS[reservoir]
        self.stream = stream or Stream()
        self.components = table.specs[0].components
        self.res_components = table.specs[0].res_components if self.components == 2 else 1
        n = table.total_points * self.components
        nr = table.total_points * self.res_components
        c, r = self.cdtype, self.rdtype
        self.whole_step = bool(whole_step)
        if self.whole_step:
            if self.components == 2:
                # the spinor whole-step kernel: single precision, periodic grids, 64 x 16 tiles
                if not spinor_whole_step_applies(table, precision):
                    raise ValueError("whole_step runs spinor fields in single precision on periodic grids "
                                     "with tile_h 16")
            elif table.tile_h not in (16, 32, 64) or (table.tile_h == 64 and precision != "single"):
                raise ValueError("whole_step needs tile_h 16 or 32 (or 64 in single precision)")
        elif table.tile_h not in (8, 16, 32):
            raise ValueError("the stage kernels need tile_h 8, 16 or 32")
        # row strips: the whole-step kernel with one lane per column and rows in registers
        self.strip_h = None if strip_h is None else int(strip_h)
        if self.strip_h is not None:
            if not self.whole_step:
                raise ValueError("row strips run the whole-step kernel (whole_step=True)")
            if self.strip_h not in (8, 16, 24, 32, 40) or (self.strip_h > 24 and precision != "single"):
                raise ValueError("strip_h must be 8, 16 or 24 (or 32 and 40 in single precision)")
            # the strip kernel addresses a member with 32-bit byte offsets
            if any(s.nx * s.ny > STRIP_MAX_POINTS for s in table.specs):
                raise ValueError(f"row strips take members of at most {STRIP_MAX_POINTS} points")
        self.psi = [DeviceArray.empty((n,), c), DeviceArray.empty((n,), c)]
        stage_bufs = 0 if self.whole_step else 3
        self.k = [DeviceArray.empty((n,), c) for _ in range(stage_bufs)]
        self.kres = [[DeviceArray.empty((nr,), r) for _ in range(self.nres)] for _ in range(stage_bufs)]
        self.res = [DeviceArray.zeros((nr,), r) for _ in range(self.nres)]
        self._res_alt = [DeviceArray.zeros((nr,), r) for _ in range(self.nres)] if self.whole_step else None
        self.step_base = DeviceArray.zeros((1,), np.uint64)
        self.potential = None      # packed complex potential (DeviceArray) or None; stage kernels only
        # active tile list: a map of fixed capacity (every tile of the table) whose leading
        # ``num_tiles`` entries run; the count also lives on the device, where the kernels read
        # it, so retiring members rewrites both in place and captured graphs stay valid
        launch_map = table.tile_map_host if self.strip_h is None else table.strip_pairs(self.strip_h)
        self.capacity = int(launch_map.shape[0])
        self.tile_map = DeviceArray.from_numpy(launch_map.reshape(-1))
        self.num_tiles = self.capacity
        # workgroups per launch.  Kernels read the active count from the device, so a graph
        # sized for more tiles stays correct, but each surplus block still costs a dispatch with
        # its LDS: a launch sized to the next power of two above the count ran 15-18% slower
        # over the 256-member sweep than one sized exactly, re-captured at each retirement
        # (a capture of ~20 launches, no allocation, after the stream is idle)
        self.launch_tiles = self.capacity
        self.active = DeviceArray.from_numpy(np.asarray([self.num_tiles], dtype=np.int32))
        self._graphs: dict = {}
        self._args = N.Rk4Args()

    def set_active_members(self, members) -> None:
        """Restrict launches to the tiles (or strips) of ``members``."""
        tab = self.table
        self.set_active_tiles(tab.tiles_of(members) if self.strip_h is None else tab.strip_pairs(self.strip_h, members))

    def set_active_tiles(self, pairs: np.ndarray) -> None:
        """Restrict launches to the listed (member, local tile) pairs: the pairs are written into
        the head of the existing map and the count into device memory (no allocation, no new
        graph).  The upload waits for the stream, so launches already queued finish first."""
        pairs = np.asarray(pairs, dtype=np.int32).reshape(-1)
        count = pairs.size // 2
        if count > self.capacity:
            raise ValueError("more active tiles than the member table holds")
        if count:
            self.tile_map.view_range(0, (pairs.size,)).upload(pairs, self.stream)
        self.active.upload(np.asarray([count], dtype=np.int32), self.stream)
        self.num_tiles = count
        if count != self.launch_tiles:
            # the uploads above waited for the stream: no captured graph is in flight
            self.launch_tiles = count
            self._graphs.clear()

    def launch_step(self, psi_in: DeviceArray, psi_out: DeviceArray, coef_ptrs, noise: bool, step: int,
                    step_base_ptr=None, stages=(0, 1, 2, 3), res_in=None, res_out=None) -> None:
        """One step from ``psi_in`` into ``psi_out``.  The stage path updates ``self.res``
        in place; the whole-step path reads ``res_in`` and writes ``res_out``."""
        if self.num_tiles == 0:
            return
        tab = self.table
        a = self._args
        a.precision = self.prec_code
        a.reservoir = N.RESERVOIR_KIND[self.reservoir]
        a.num_tiles = self.launch_tiles
        a.active_tiles = self.active.ptr
        a.tile_y = tab.tile_h if self.strip_h is None else self.strip_h
        a.strips = 0 if self.strip_h is None else 1
        a.no_eta = 1 if all(r.eta == 0.0 for r in tab.rows) else 0
        a.five_point = 1 if all(r.laplacian == N.LAPLACIAN_KIND["five-point"] for r in tab.rows) else 0
        a.members = tab.members
        a.components = self.components
        a.res_components = self.res_components
        a.dt = float(self._dt)
        a.table = tab.table_ptr
        a.tile_map = self.tile_map.ptr
        a.psi0 = psi_in.ptr
        a.pump = self.pump.ptr
        a.potential = self.potential.ptr if self.potential is not None else None
        lib = N.lib()
        if self.whole_step:
            for f in range(self.nres):
                a.res0[f] = res_in[f].ptr
                a.res_out[f] = res_out[f].ptr
            a.psi_out = psi_out.ptr
            a.coef = coef_ptrs[0]          # the 4 stage rows of this step are contiguous
            a.noise = 1 if noise else 0
            a.step = int(step)
            a.step_base = step_base_ptr
            N.check(lib.pgpe_rk4_step(C.byref(a), self.stream.handle), "pgpe_rk4_step")
            return
        for f in range(self.nres):
            a.res0[f] = self.res[f].ptr
        for s in stages:
            a.stage = s
            a.coef = coef_ptrs[s]
            if s > 0:
                a.kprev = self.k[s - 1].ptr
                for f in range(self.nres):
                    a.kres_prev[f] = self.kres[s - 1][f].ptr
            if s < 3:
                a.kout = self.k[s].ptr
                a.noise = 0
                for f in range(self.nres):
                    a.kres_out[f] = self.kres[s][f].ptr
            else:
                a.k1, a.k2 = self.k[0].ptr, self.k[1].ptr
                a.psi_out = psi_out.ptr
                for f in range(self.nres):
                    a.kres1[f] = self.kres[0][f].ptr
                    a.kres2[f] = self.kres[1][f].ptr
                    a.res_out[f] = self.res[f].ptr
                a.noise = 1 if noise else 0
                a.step = int(step)
                a.step_base = step_base_ptr
            N.check(lib.pgpe_rk4_stage(C.byref(a), self.stream.handle), "pgpe_rk4_stage")

    def set_dt(self, dt: float) -> None:
        self._dt = float(dt)
        self._graphs.clear()

    def coef_ptrs(self, coef: DeviceArray, index: int) -> list[int]:
        item = np.dtype(self.rdtype).itemsize
        m = self.table.members
        return [coef.ptr + ((index * 4 + s) * m) * item for s in range(4)]

    MAX_GRAPH_STEPS = 1000

    def run_steps(self, count: int, coef: DeviceArray, const: bool, noise: bool, step0: int,
                  graphs: bool = True) -> None:
        """Advance ``count`` steps from ``psi[0]``; the result ends in ``psi[0]``.

        With ``graphs`` the even part of ``count`` replays one captured graph
        (coefficients and the noise step base are read from device buffers).
        """
        if count > 0 and not const and coef.size < count * 4 * self.table.members:
            raise ValueError("coefficient buffer holds fewer steps than requested")
        done = 0
        while done < count:
            rem = count - done
            if graphs and rem >= 2:
                chunk = rem - rem % 2
                if const:   # constant coefficients: replay one graph of at most MAX_GRAPH_STEPS steps
                    chunk = min(chunk, self.MAX_GRAPH_STEPS)
                p0, p1 = self.psi
                key = (chunk, const, noise, self.launch_tiles, p0.ptr, coef.ptr,
                       self.res[0].ptr if self.whole_step and self.nres else 0,
                       self.potential.ptr if self.potential is not None else 0)
                g = self._graphs.get(key)
                if g is None:
                    r0, r1 = self.res, self._res_alt
                    def body():
                        for k in range(chunk):
                            src, dst = (p0, p1) if k % 2 == 0 else (p1, p0)
                            rsrc, rdst = (r0, r1) if k % 2 == 0 else (r1, r0)
                            self.launch_step(src, dst, self.coef_ptrs(coef, 0 if const else k), noise,
                                             k, self.step_base.ptr, res_in=rsrc, res_out=rdst)
                    self.stream.synchronize()
                    g = Graph.capture(self.stream, body)
                    if len(self._graphs) >= 8:
                        self._graphs.pop(next(iter(self._graphs)))
                    self._graphs[key] = g
                self.step_base.upload(np.asarray([step0 + done], dtype=np.uint64), self.stream)
                g.launch(self.stream)
                done += chunk
            else:
                self.launch_step(self.psi[0], self.psi[1], self.coef_ptrs(coef, 0 if const else done), noise,
                                 step0 + done, None, res_in=self.res, res_out=self._res_alt)
                self.psi.reverse()
                if self.whole_step:
                    self.res, self._res_alt = self._res_alt, self.res
                done += 1
