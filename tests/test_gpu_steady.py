"""The device steady-state test (simulation_controller.py:710-734) against the oracle.

* ``pgpe_density_delta``: the three per-member FP64 sums (sum rho, sum (rho - old)^2, sum old^2)
  on random uneven-grid batches, scalar and spinor, full and partial tile maps, equal the
  NumPy float64 sums, and the (relative density, relative norm) changes built from them equal
  the oracle's ``steady_metrics``.
* ``PackedSweep``: the stop times of a four-member uneven-grid sweep, with converged members
  retired from the tile map, equal those of an oracle loop that steps each member with
  ``rk4_fdm_step`` and applies ``steady_metrics`` every ``check_every``; the retired members'
  final fields are the oracle's fields at their stop times.
"""
import numpy as np
import pytest

from oracle import gpe_oracle as O

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (100, 37), (33, 130), (128, 128)]


def _table(sizes, components=1, tile_h=16):
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.layout import MemberSpec, MemberTable
    specs = [MemberSpec(nx, ny, 0.5, 0.7, PhysicsConstants(), components=components,
                        res_components=2 if components == 2 else 1) for nx, ny in sizes]
    return MemberTable(specs, tile_h)


def _host_sums(psi, old):
    rho = (psi.real * psi.real + psi.imag * psi.imag).sum(axis=0)
    return np.array([rho.astype(np.float64).sum(), ((rho.astype(np.float64) - old) ** 2).sum(),
                     (old.astype(np.float64) ** 2).sum()]), rho


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("components", [1, 2])
def test_density_sums_equal_host_float64_sums(precision, components):
    from polarism_amd import _native as N
    from polarism_amd.device import DeviceArray, init
    init()
    rng = np.random.default_rng(7 + components)
    c = np.complex64 if precision == "single" else np.complex128
    r = np.float32 if precision == "single" else np.float64
    tab = _table(SIZES, components)
    fields, olds = [], []
    for nx, ny in SIZES:
        fields.append((rng.standard_normal((components, ny, nx)) + 1j * rng.standard_normal((components, ny, nx))).astype(c))
        olds.append(np.abs(rng.standard_normal((ny, nx))).astype(r))
    psi_dev = DeviceArray.from_numpy(np.concatenate([f.reshape(-1) for f in fields]))
    old_dev = DeviceArray.from_numpy(np.concatenate([o.reshape(-1) for o in olds]))
    partials = DeviceArray.zeros((3 * tab.num_tiles,), np.float64)
    sums = DeviceArray.zeros((3 * tab.members,), np.float64)
    prec = 0 if precision == "single" else 1
    N.call("pgpe_density_delta", psi_dev.ptr, old_dev.ptr, prec, components, tab.table_ptr, tab.tile_map_ptr,
           tab.num_tiles, tab.members, partials.ptr, sums.ptr, 1, None)
    dev = sums.get().reshape(-1, 3)
    new_old = old_dev.get()
    cell = 0.5 * 0.7
    for m, (f, o) in enumerate(zip(fields, olds)):
        host, rho = _host_sums(f, o)
        np.testing.assert_allclose(dev[m], host, rtol=1e-12)
        off = int(tab.offsets[m])
        np.testing.assert_array_equal(new_old[off:off + o.size], rho.reshape(-1))   # the reference density moves on
        _, norms, rel_d, rel_n = O.steady_metrics(f[None], o.astype(np.float64), cell)
        dev_rel_d = np.sqrt(dev[m, 1]) / max(np.sqrt(dev[m, 2]), 1e-30)
        dev_norm = dev[m, 0] * cell
        # the device sums in FP64; the oracle forms the density difference in the field precision
        tol = 1e-12 if precision == "double" else 1e-6
        assert abs(dev_rel_d - rel_d[0]) <= tol * rel_d[0]
        assert abs(dev_norm - norms[0]) <= tol * norms[0]
        old_norm = o.astype(np.float64).sum() * cell
        assert abs(abs(dev_norm - old_norm) / old_norm - rel_n[0]) <= 1e3 * tol * max(rel_n[0], 1e-300)
    # a partial tile map (members 1 and 3 retired): the running members' sums are unchanged
    keep = tab.tiles_of([0, 2])
    tmap = DeviceArray.from_numpy(keep)
    old_dev.upload(np.concatenate([o.reshape(-1) for o in olds]))
    N.call("pgpe_density_delta", psi_dev.ptr, old_dev.ptr, prec, components, tab.table_ptr, tmap.ptr,
           int(keep.shape[0]), tab.members, partials.ptr, sums.ptr, 0, None)
    dev2 = sums.get().reshape(-1, 3)
    for m in (0, 2):
        np.testing.assert_allclose(dev2[m], dev[m], rtol=1e-15)


def _uniform_steady(p, P0):
    """Single reservoir, uniform pump above threshold: nR = gamma_C / R, |psi|^2 = (P0 / nR - gamma_R) / R."""
    nR = p.gamma_C / p.R
    return nR, (P0 / nR - p.gamma_R) / p.R


def _oracle_stop_time(m, p, dt, steps_per_check, rtol, max_checks):
    g = O.Grid(m.nx, m.ny, m.lx, m.ly)
    psi = m.psi0.reshape(1, 1, m.ny, m.nx).astype(np.complex128)
    res = (m.res0[0].reshape(1, 1, m.ny, m.nx).astype(np.float64),)
    pump = (m.P0 * m.envelope).reshape(1, 1, m.ny, m.nx)
    old = None
    cell = g.dx * g.dy
    step = 0
    for _ in range(max_checks):
        for _ in range(steps_per_check):
            psi, res = O.rk4_fdm_step("single", p, g, psi, res, lambda t: pump, step * dt, dt)
            step += 1
        density, norms, rel_d, rel_n = O.steady_metrics(psi, old, cell)
        if old is not None:
            reached = norms[0] < 1e-30 or (rel_n[0] < rtol and rel_d[0] < rtol)
            if reached:
                return step * dt, psi[0, 0]
        old = density
    return np.nan, psi[0, 0]


@pytest.mark.parametrize("form,tile_h", [("stage", 16), ("step", 16)])
def test_packed_sweep_stop_times_equal_oracle_loop(form, tile_h):
    """Four members of three grid shapes start from perturbed uniform steady states and relax;
    each reaches the criterion at a different check.  Stop times are exact multiples of
    check_every, so they must agree exactly."""
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.sweep import PackedSweep, SweepMember
    p_cfg = PhysicsConstants(kinetic_relaxation_eta=0.3)     # damps the perturbation's k != 0 modes
    p = O.Physics(kinetic_relaxation_eta=0.3)
    dt, check_every, rtol = 0.05, 1.0, 1e-3
    members = []
    shapes = [(32, 32, 24.0, 24.0), (48, 32, 36.0, 24.0), (32, 32, 24.0, 24.0), (64, 48, 48.0, 36.0)]
    perturb = [0.01, 0.3, 0.1, 0.05]
    P0s = [0.03, 0.05, 0.04, 0.06]
    for i, ((nx, ny, lx, ly), eps, P0) in enumerate(zip(shapes, perturb, P0s)):
        nR, rho = _uniform_steady(p, P0)
        g = O.Grid(nx, ny, lx, ly)
        wave = np.cos(2 * np.pi * g.X / lx) * np.cos(2 * np.pi * 2 * g.Y / ly)
        psi0 = np.sqrt(rho) * (1 + eps * wave) * np.exp(0.3j * wave)
        res0 = (np.full((ny, nx), nR) * (1 - 0.5 * eps * wave),)
        members.append(SweepMember(nx=nx, ny=ny, lx=lx, ly=ly, physics=p_cfg, envelope=np.ones((ny, nx)), P0=P0,
                                   seed=i, psi0=psi0, res0=res0))
    sw = PackedSweep(members, reservoir="single", precision="double", dt=dt, check_every=check_every, rtol=rtol,
                     max_time=40.0, tile_h=tile_h, form=form, retire=True, graphs=True)
    out = sw.run()
    steps = int(round(check_every / dt))
    expect = []
    for i, m in enumerate(members):
        t_stop, psi_stop = _oracle_stop_time(m, p, dt, steps, rtol, 40)
        expect.append(t_stop)
        if np.isfinite(t_stop):
            assert O.rel_l2(out.final_psi[i], psi_stop) <= 1e-10
    expect = np.asarray(expect)
    print(f"{form}: stop times HIP {out.stop_times} oracle {expect}", flush=True)
    assert np.all(np.isfinite(expect)) and len(set(expect.tolist())) == 4
    np.testing.assert_array_equal(out.stop_times, expect)
