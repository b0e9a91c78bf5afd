"""RK4 launch engine over a packed member table.

Owns the stage buffers of one batch (uniform or uneven grids) and issues the
four fused ``pgpe_rk4_stage`` launches of a step.  Used by the
``rk4-fdm-hip`` solver and by the packed sweep driver.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .device import DeviceArray, Graph, Stream
from .layout import MemberTable


def real_dtype(precision):
    return np.float32 if precision == "single" else np.float64


def complex_dtype(precision):
    return np.complex64 if precision == "single" else np.complex128


class Rk4Engine:
    def __init__(self, table: MemberTable, pump: DeviceArray, precision: str, reservoir: str,
                 stream: Stream | None = None):
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
        self.psi = [DeviceArray.empty((n,), c), DeviceArray.empty((n,), c)]
        self.k = [DeviceArray.empty((n,), c) for _ in range(3)]
        self.kres = [[DeviceArray.empty((nr,), r) for _ in range(self.nres)] for _ in range(3)]
        self.res = [DeviceArray.zeros((nr,), r) for _ in range(self.nres)]
        self.step_base = DeviceArray.zeros((1,), np.uint64)
        # active tile list: starts as every tile; the sweep driver may retire members
        self.tile_map = table.tile_map_dev
        self.num_tiles = table.num_tiles
        self._graphs: dict = {}
        self._args = N.Rk4Args()

    def set_active_tiles(self, pairs: np.ndarray) -> None:
        """Restrict launches to the listed (member, local tile) pairs."""
        pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
        if pairs.shape[0] == 0:
            self.num_tiles = 0
            return
        self.tile_map = DeviceArray.from_numpy(pairs)
        self.num_tiles = int(pairs.shape[0])
        self._graphs.clear()

    def launch_step(self, psi_in: DeviceArray, psi_out: DeviceArray, coef_ptrs, noise: bool, step: int,
                    step_base_ptr=None, stages=(0, 1, 2, 3)) -> None:
        if self.num_tiles == 0:
            return
        tab = self.table
        a = self._args
        a.precision = self.prec_code
        a.reservoir = N.RESERVOIR_KIND[self.reservoir]
        a.num_tiles = self.num_tiles
        a.tile_y = tab.tile_h
        a.members = tab.members
        a.components = self.components
        a.res_components = self.res_components
        a.dt = float(self._dt)
        a.table = tab.table_ptr
        a.tile_map = self.tile_map.ptr
        a.psi0 = psi_in.ptr
        a.pump = self.pump.ptr
        for f in range(self.nres):
            a.res0[f] = self.res[f].ptr
        lib = N.lib()
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
                p0, p1 = self.psi
                key = (chunk, const, noise, self.num_tiles, self.tile_map.ptr, p0.ptr, coef.ptr)
                g = self._graphs.get(key)
                if g is None:
                    def body():
                        for k in range(chunk):
                            src, dst = (p0, p1) if k % 2 == 0 else (p1, p0)
                            self.launch_step(src, dst, self.coef_ptrs(coef, 0 if const else k), noise,
                                             k, self.step_base.ptr)
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
                                 step0 + done, None)
                self.psi.reverse()
                done += 1
