"""Packed member layout and the device member table.

Members of a batch (equal or uneven grids) are packed back to back in every
per-point device array; member ``m`` starts at ``offsets[m]`` grid points.
Each member is cut into 64 x ``tile_h`` tiles; ``tile_map[t] = (member,
local tile)`` lists the tiles one launch covers, so a subset (members that
have converged removed) is just a shorter list.  One launch of any stencil kernel covers every
tile of every member, which is how an uneven-grid sweep runs as one launch
(an addition; the reference runs each grid shape as its own batch,
sweep_config.py:18 and sweep.py:128-138).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .device import DeviceArray
from .reservoir import reservoir_rates

TILE_W = 64
STRIP_W = 56      # output columns of a row strip: a 64-lane wavefront less a 4-column halo each side


@dataclass
class MemberSpec:
    """Everything the device needs to know about one member."""
    nx: int
    ny: int
    dx: float
    dy: float
    physics: object            # attribute access: hbar, m_eff, g_C, ...
    reservoir: str = "single"
    components: int = 1
    res_components: int = 1
    pump_off: int = 0          # points; members sharing one envelope use the same offset
    seed: int = 0
    noise_strength: float = 0.0
    dt: float = 1e-3
    circular_polarization: float = 0.0
    laplacian: str = "five-point"     # or "isotropic-9pt" (dx == dy)
    boundary: str = "periodic"        # or "closed-interval" (mirrored Neumann edges)
    pot_off: int = 0                  # points into the packed complex potential, if any


class MemberTable:
    def __init__(self, specs: list[MemberSpec], tile_h: int = 16):
        if tile_h not in (8, 16, 32, 64):
            raise ValueError("tile_h must be 8, 16, 32 or 64")
        self.specs = list(specs)
        self.tile_h = tile_h
        self.members = len(self.specs)
        rows = (N.Member * self.members)()
        offsets = []
        tile_map = []
        off = 0
        tile = 0
        for m, s in enumerate(self.specs):
            p = s.physics
            get = lambda name, p=p: float(getattr(p, name))
            row = rows[m]
            row.off = off
            row.pump_off = s.pump_off
            row.nx, row.ny = s.nx, s.ny
            row.tiles_x = -(-s.nx // TILE_W)
            row.tiles_y = -(-s.ny // tile_h)
            row.tile_begin = tile
            row.components = s.components
            row.tile_h = tile_h
            row.res_components = s.res_components
            row.inv_dx2 = 1.0 / (s.dx ** 2)
            row.inv_dy2 = 1.0 / (s.dy ** 2)
            row.hbar_over_2m = get("hbar") / (2.0 * get("m_eff"))
            row.inv_hbar = 1.0 / get("hbar")
            row.kin = -(get("hbar") ** 2) / (2 * get("m_eff"))
            row.hbar, row.m_eff = get("hbar"), get("m_eff")
            row.g_C, row.g_R, row.g_I = get("g_C"), get("g_R"), get("g_I")
            row.R, row.gamma_C = get("R"), get("gamma_C")
            row.eta = get("kinetic_relaxation_eta")
            row.res_a, row.res_b, row.res_c, row.res_d = reservoir_rates(s.reservoir, get)
            row.alpha_2, row.zeeman, row.te_tm = get("alpha_2"), get("zeeman"), get("te_tm")
            row.inv_4dxdy = 1.0 / (4 * s.dx * s.dy)
            row.laplacian = N.LAPLACIAN_KIND[s.laplacian]
            row.boundary = N.BOUNDARY_KIND[s.boundary]
            row.pot_off = s.pot_off
            if s.laplacian == "isotropic-9pt" and abs(s.dx - s.dy) > 1e-8 * max(s.dx, s.dy):
                raise ValueError(f"laplacian 'isotropic-9pt' requires dx == dy, got dx={s.dx}, dy={s.dy}")
            if s.components == 2 and s.laplacian != "five-point":
                raise ValueError("spinor members run the five-point stencil")
            row.noise_gamma = get("gamma_C")
            row.noise_R = get("R")
            row.noise_strength = float(s.noise_strength)
            row.noise_dt_over_4v = s.dt / (4.0 * (s.dy * s.dx))
            row.seed = int(s.seed) & 0xFFFFFFFFFFFFFFFF
            row.pump_w_plus = (1 + s.circular_polarization) / 2
            row.pump_w_minus = (1 - s.circular_polarization) / 2
            n_tiles = row.tiles_x * row.tiles_y
            tile_map.extend((m, j) for j in range(n_tiles))
            offsets.append(off)
            off += s.nx * s.ny
            tile += n_tiles
        self.rows = rows
        self.offsets = np.asarray(offsets, dtype=np.int64)
        self.total_points = off
        self.num_tiles = tile
        self.tile_map_host = np.asarray(tile_map, dtype=np.int32).reshape(-1, 2)
        raw = np.frombuffer(bytes(rows), dtype=np.uint8)
        self.table_dev = DeviceArray.from_numpy(raw)
        self.tile_map_dev = DeviceArray.from_numpy(self.tile_map_host)

    @property
    def table_ptr(self) -> int:
        return self.table_dev.ptr

    @property
    def tile_map_ptr(self) -> int:
        return self.tile_map_dev.ptr

    def tiles_of(self, members) -> np.ndarray:
        """(member, local) pairs of the given members, in table order."""
        keep = np.isin(self.tile_map_host[:, 0], np.asarray(list(members), dtype=np.int32))
        return self.tile_map_host[keep]

    def strip_pairs(self, strip_h: int, members=None) -> np.ndarray:
        """(member, strip) pairs of the row-strip whole-step launch: member m is cut into
        ceil(nx / STRIP_W) x ceil(ny / strip_h) strips of STRIP_W output columns, listed row by
        row, each as (row << 16) | column; ``members`` (default all) selects whose strips, in
        table order."""
        keep = range(self.members) if members is None else sorted(int(m) for m in members)
        out = []
        for m in keep:
            s = self.specs[m]
            cols, rows = -(-s.nx // STRIP_W), -(-s.ny // int(strip_h))
            r, c = np.divmod(np.arange(rows * cols, dtype=np.int64), cols)
            local = ((r << 16) | c).astype(np.int32)
            out.append(np.stack([np.full(rows * cols, m, dtype=np.int32), local], axis=1))
        return np.concatenate(out) if out else np.zeros((0, 2), dtype=np.int32)

    def member_points(self, m: int) -> int:
        s = self.specs[m]
        return s.nx * s.ny

    def set_pot_offsets(self, offsets) -> None:
        """Point member m at offsets[m] of the packed potential and re-upload the table."""
        for m, o in enumerate(offsets):
            self.specs[m].pot_off = int(o)
            self.rows[m].pot_off = int(o)
        self.refresh()

    def refresh(self) -> None:
        """Re-upload the table after host-side edits of ``rows``."""
        raw = np.frombuffer(bytes(self.rows), dtype=np.uint8)
        self.table_dev.upload(raw)
