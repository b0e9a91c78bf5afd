"""The reference-side plugin adapter (polarism_amd.plugin_adapter) driven through stand-ins that
carry the reference's AbstractSolver / StepReport / state / reservoir signatures
(abstract_solver.py:60-101, simulation_state.py:92-192).  The HIP controller is replaced by a
CPU stand-in that steps with the oracle, so the test runs without a GPU and checks the adapter's
own duties: configuration and member translation, step / t / t0 bookkeeping, noise applied once,
and no aliasing between the reference's arrays and the HIP side's."""
from __future__ import annotations

import dataclasses
from collections import namedtuple

import numpy as np

from oracle import gpe_oracle as O
from polarism_amd.plugin_adapter import config_mapping, make_hip_solver, member_overrides
from tests.helpers import oracle_physics, quickstart_config

RefReport = namedtuple("StepReport", "accepted error dt_next")


class RefAbstractSolver:
    """The reference base class's behaviour (abstract_solver.py:71-101), fixed steps only."""
    injects_noise = False

    def __init__(self, config, grid, model):
        self.config = config
        self.physics = config.physics
        self.batch_physics = model.physics

    def advance(self, model, state, n_steps):
        for _ in range(n_steps):
            completed = state.step
            report = self.step(model, state, self.config.solver.dt)
            if report.accepted and model.noise is not None and not self.injects_noise:
                model.noise.apply_to_state(state, model.reservoir, completed)


class RefNoise:
    def __init__(self):
        self.applied = 0
        self.seeds = np.asarray([5])

    def apply_to_state(self, state, reservoir, step):
        self.applied += 1


class RefReservoir:
    def __init__(self, fields):
        self.fields = tuple(fields)

    def get_state(self):
        return self.fields

    def set_state(self, fields):
        self.fields = tuple(np.maximum(f, 0) for f in fields)


@dataclasses.dataclass
class RefState:
    field: np.ndarray
    t: float = 0.0
    t0: float = 0.0
    step: int = 0


class RefModel:
    def __init__(self, reservoir, noise, host=None):
        self.reservoir = reservoir
        self.noise = noise
        self.physics = type("BatchPhysics", (), {"host": host or {}, "size": None})()


# ---------------------------------------------------------------- the HIP side, restated on the CPU
class _Dev:
    """A device array stand-in: set from and read back as copies."""
    def __init__(self, a):
        self._a = np.array(a, copy=True)

    def get(self):
        return np.array(self._a, copy=True)


class _HipState:
    def __init__(self):
        self._field, self.step, self.t, self.t0 = None, 0, 0.0, 0.0

    @property
    def field(self):
        return self._field

    @field.setter
    def field(self, value):
        self._field = _Dev(value)


class _HipReservoir:
    def set_state(self, fields):
        self._fields = tuple(_Dev(f) for f in fields)

    def get_state(self):
        return self._fields


class _OracleSolver:
    """Fixed rk4-fdm steps with the oracle: the semantics the HIP solver has (tested on the GPU)."""
    def __init__(self, cfg, outer):
        self.cfg, self.outer = cfg, outer

    def _one(self, state, reservoir):
        cfg = self.cfg
        g = O.Grid(cfg.grid.nx, cfg.grid.ny, cfg.grid.lx, cfg.grid.ly)
        env = O.gaussian_profile(g, cfg.laser.sigma_space)
        psi = state.field.get().reshape(1, 1, *g.shape)
        res = tuple(r.get().reshape(1, 1, *g.shape) for r in reservoir.get_state())
        psi, res = O.rk4_fdm_step("single", oracle_physics(cfg), g, psi, res, lambda t: env * cfg.laser.P0,
                                  state.t, cfg.solver.dt)
        state.field = psi
        reservoir.set_state(res)

    def step(self, model, state, dt):
        """One step without noise, as the HIP solvers' step()."""
        self._one(state, self.outer.reservoir)
        state.step += 1
        state.t = state.t0 + state.step * self.cfg.solver.dt
        return RefReport(True, None, None)

    def advance(self, model, state, n):
        """Steps with the Wigner increment injected inside (counted)."""
        for _ in range(n):
            self.step(model, state, self.cfg.solver.dt)
            self.outer.noise_injected += 1


class _HipController:
    def __init__(self, cfg, members):
        self.cfg, self.members = cfg, members
        self.state, self.reservoir, self.model = _HipState(), _HipReservoir(), None
        self.noise_injected = 0
        self.solver = _OracleSolver(cfg, self)


# ---------------------------------------------------------------- tests
def _setup():
    cfg = quickstart_config(grid__nx=16, grid__ny=16, grid__lx=16.0, grid__ly=16.0, solver__dt=0.01,
                            laser__laser_type="continuous-gaussian", laser__P0=0.05, laser__sigma_space=4.0)
    made = []

    def factory(c, members):
        made.append(_HipController(c, members))
        return made[-1]

    Adapter = make_hip_solver("rk4-fdm-hip", base=RefAbstractSolver, report_cls=RefReport, to_host=np.asarray,
                              to_array=np.asarray, controller_factory=factory)
    rng = np.random.default_rng(2)
    psi0 = 0.05 * (rng.standard_normal((1, 1, 16, 16)) + 1j * rng.standard_normal((1, 1, 16, 16)))
    model = RefModel(RefReservoir((np.full((1, 1, 16, 16), 0.3),)), RefNoise())
    state = RefState(psi0.copy(), t=0.5, t0=0.5)
    solver = Adapter(cfg, None, model)
    return cfg, solver, model, state, made[0], psi0


def test_adapter_advance_bookkeeping_noise_once_and_no_aliasing():
    cfg, solver, model, state, hip, psi0 = _setup()
    assert solver.injects_noise and type(solver).__name__ == "Hip_rk4_fdm_hip"
    solver.advance(model, state, 5)
    assert state.step == 5 and state.t == state.t0 + 5 * cfg.solver.dt
    assert model.noise.applied == 0 and hip.noise_injected == 5   # once per step, by the HIP side
    # the same five steps with the oracle directly
    g = O.Grid(16, 16, 16.0, 16.0)
    env = O.gaussian_profile(g, 4.0)
    psi, res = psi0, (np.full((1, 1, 16, 16), 0.3),)
    for s in range(5):
        psi, res = O.rk4_fdm_step("single", oracle_physics(cfg), g, psi, res, lambda t: env * 0.05, 0.5 + s * 0.01,
                                  0.01)
    np.testing.assert_allclose(state.field, psi, rtol=0, atol=1e-15)
    np.testing.assert_allclose(model.reservoir.get_state()[0], res[0], rtol=0, atol=1e-15)
    # no aliasing: writing into the reference's arrays leaves the HIP side untouched
    before = hip.state.field.get()
    state.field[...] = 0
    model.reservoir.get_state()[0][...] = 0
    np.testing.assert_array_equal(hip.state.field.get(), before)
    assert np.any(hip.reservoir.get_state()[0].get() != 0)


def test_adapter_step_moves_one_tick_and_reports():
    cfg, solver, model, state, hip, _ = _setup()
    report = solver.step(model, state, cfg.solver.dt)
    assert report.accepted and report.error is None
    assert state.step == 1 and state.t == state.t0 + cfg.solver.dt
    assert model.noise.applied == 0 and hip.noise_injected == 1
    # the reference's own loop around step() (abstract_solver.py:97-101): still exactly one
    # increment per step, from the HIP side, none from the reference
    RefAbstractSolver.advance(solver, model, state, 3)
    assert state.step == 4 and state.t == state.t0 + 4 * cfg.solver.dt
    assert model.noise.applied == 0 and hip.noise_injected == 4


def test_config_mapping_and_members():
    cfg = quickstart_config()
    ignored = []
    data = config_mapping(cfg, ignored)
    assert data["grid"]["nx"] == 128 and not ignored
    data = dataclasses.asdict(cfg)
    data["grid"]["future_key"] = 1
    data["newsection"] = {}
    out = config_mapping(data, ignored)
    assert "future_key" not in out["grid"] and set(ignored) == {"grid.future_key", "newsection"}
    model = RefModel(None, RefNoise(), host={"g_C": np.asarray([1e-3])})
    model.physics.size = 1
    assert member_overrides(model) == [{"physics.g_C": 1e-3, "noise.seed": 5}]
    model.physics.host = {}
    assert member_overrides(model) is None
