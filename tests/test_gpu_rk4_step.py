"""Whole-step RK4 kernel (pgpe_rk4_step), on LDS tiles and on register row strips, against the
four stage launches, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (96, 80), (200, 72), (128, 128), (8, 40)]


def _engines(sizes, reservoir, precision, tile_h, noise_strength=0.0, laplacian="five-point",
             boundary="periodic", strip_h=None):
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.device import DeviceArray, init
    from polarism_amd.engine import Rk4Engine, real_dtype
    from polarism_amd.layout import MemberSpec, MemberTable
    init()
    p = PhysicsConstants()
    p.g_I = 0.0005
    p.kinetic_relaxation_eta = 0.02
    rng = np.random.default_rng(7)
    specs, envs, off = [], [], 0
    for i, (nx, ny) in enumerate(sizes):
        specs.append(MemberSpec(nx=nx, ny=ny, dx=0.5, dy=0.5 if laplacian == "isotropic-9pt" else 0.6,
                                physics=p, reservoir=reservoir, pump_off=off, seed=31 + i,
                                noise_strength=noise_strength, dt=0.05, laplacian=laplacian, boundary=boundary))
        envs.append(rng.uniform(0.0, 1.0, nx * ny))
        off += nx * ny
    r = real_dtype(precision)
    pump = DeviceArray.from_numpy(np.concatenate(envs).astype(r))
    coef = DeviceArray.from_numpy(np.tile(rng.uniform(0.02, 0.08, len(sizes)), (4, 1)).astype(r))
    stage_tab = MemberTable(specs, 16 if tile_h == 64 else tile_h)
    step_tab = MemberTable(specs, tile_h)
    e_stage = Rk4Engine(stage_tab, pump, precision, reservoir)
    e_step = Rk4Engine(step_tab, pump, precision, reservoir, whole_step=True, strip_h=strip_h)
    total = stage_tab.total_points
    psi = (0.3 * (rng.standard_normal(total) + 1j * rng.standard_normal(total))).astype(e_stage.cdtype)
    res = [rng.uniform(0.0, 3.0, total).astype(r) for _ in range(e_stage.nres)]
    for e in (e_stage, e_step):
        e.set_dt(0.05)
        e.psi[0].upload(psi)
        for f in range(e.nres):
            e.res[f].upload(res[f])
    return e_stage, e_step, coef


def _compare(e_stage, e_step, rtol=0.0):
    """Bitwise (rtol 0), or to rounding: on a closed interval the whole-step kernel computes its
    mirrored halo points itself, whose stencil sums then run in mirrored order."""
    e_stage.stream.synchronize()
    e_step.stream.synchronize()
    a, b = e_stage.psi[0].get(), e_step.psi[0].get()
    assert np.all(np.isfinite(a))
    pairs = [(a, b)] + [(e_stage.res[f].get(), e_step.res[f].get()) for f in range(e_stage.nres)]
    for x, y in pairs:
        if rtol == 0.0:
            np.testing.assert_array_equal(x, y)
        else:
            np.testing.assert_allclose(y, x, rtol=rtol, atol=rtol * np.max(np.abs(x)))


@pytest.mark.parametrize("reservoir", ["none", "single", "double", "quadratic-double"])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_whole_step_equals_stage_launches_bitwise(reservoir, precision):
    e_stage, e_step, coef = _engines(SIZES, reservoir, precision, 16)
    for e in (e_stage, e_step):
        e.run_steps(5, coef, True, False, 0, graphs=False)
    _compare(e_stage, e_step)


@pytest.mark.parametrize("tile_h,precision", [(32, "single"), (64, "single"), (32, "double")])
def test_whole_step_tile_heights_and_graph_chunks(tile_h, precision):
    """Graph-captured chunks (even) plus a trailing single step, with Wigner noise fused in."""
    e_stage, e_step, coef = _engines(SIZES, "single", precision, tile_h, noise_strength=1.0)
    for e in (e_stage, e_step):
        e.run_steps(7, coef, True, True, 3, graphs=True)
    _compare(e_stage, e_step)


@pytest.mark.parametrize("laplacian", ["five-point", "isotropic-9pt"])
@pytest.mark.parametrize("boundary", ["periodic", "closed-interval"])
def test_whole_step_stencils_and_boundaries(laplacian, boundary):
    """The 4-point halo of the whole-step kernel is exact for the mirrored closed interval too
    (bit for bit on periodic grids, to rounding on the closed interval)."""
    e_stage, e_step, coef = _engines(SIZES, "double", "double", 16, laplacian=laplacian, boundary=boundary)
    for e in (e_stage, e_step):
        e.run_steps(6, coef, True, False, 0, graphs=True)
    _compare(e_stage, e_step, rtol=0.0 if boundary == "periodic" else 1e-13)


def test_whole_step_with_retired_members():
    """A tile map covering a subset of members leaves the others untouched."""
    e_stage, e_step, coef = _engines(SIZES, "double", "single", 32)
    keep = [0, 2, 4]
    for e in (e_stage, e_step):
        e.set_active_tiles(e.table.tiles_of(np.asarray(keep)))
        e.run_steps(4, coef, True, False, 0, graphs=True)
    _compare(e_stage, e_step)


def test_whole_step_rejects_aliased_reservoir():
    from polarism_amd import _native as N
    e_stage, e_step, coef = _engines(SIZES[:1], "single", "single", 16)
    with pytest.raises(RuntimeError, match="alias"):
        e_step.launch_step(e_step.psi[0], e_step.psi[1], e_step.coef_ptrs(coef, 0), False, 0,
                           res_in=e_step.res, res_out=e_step.res)


# ---------------------------------------------------------------- row strips
@pytest.mark.parametrize("reservoir", ["none", "single", "double", "quadratic-double"])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_strips_equal_stage_launches_bitwise(reservoir, precision):
    """Sizes include a grid narrower than one strip (8 columns: the lanes past it hold wrapped
    columns) and widths that leave a partial last strip (200 = 3 x 56 + 32)."""
    e_stage, e_strip, coef = _engines(SIZES, reservoir, precision, 16, strip_h=16)
    assert e_strip.capacity == sum(-(-nx // 56) * -(-ny // 16) for nx, ny in SIZES)
    for e in (e_stage, e_strip):
        e.run_steps(5, coef, True, False, 0, graphs=False)
    _compare(e_stage, e_strip)


@pytest.mark.parametrize("strip_h,precision", [(8, "single"), (8, "double"), (24, "single"), (32, "single"), (40, "single"),
                                              (24, "double")])
def test_strip_heights_and_graph_chunks(strip_h, precision):
    """Graph-captured chunks plus a trailing single step, Wigner noise fused in."""
    e_stage, e_strip, coef = _engines(SIZES, "single", precision, 16, noise_strength=1.0, strip_h=strip_h)
    for e in (e_stage, e_strip):
        e.run_steps(7, coef, True, True, 3, graphs=True)
    _compare(e_stage, e_strip)


@pytest.mark.parametrize("laplacian", ["five-point", "isotropic-9pt"])
@pytest.mark.parametrize("boundary", ["periodic", "closed-interval"])
def test_strip_stencils_and_boundaries(laplacian, boundary):
    e_stage, e_strip, coef = _engines(SIZES, "double", "double", 16, laplacian=laplacian, boundary=boundary,
                                      strip_h=24)
    assert e_strip.capacity == sum(-(-nx // 56) * -(-ny // 24) for nx, ny in SIZES)
    for e in (e_stage, e_strip):
        e.run_steps(6, coef, True, False, 0, graphs=True)
    _compare(e_stage, e_strip, rtol=0.0 if boundary == "periodic" else 1e-13)


def test_strips_with_retired_members():
    """Retiring members rewrites the strip map in place; the retired members stay untouched."""
    e_stage, e_strip, coef = _engines(SIZES, "double", "single", 16, strip_h=24)
    for e in (e_stage, e_strip):
        e.run_steps(2, coef, True, False, 0, graphs=True)
        e.set_active_members([0, 2, 4])
        e.run_steps(4, coef, True, False, 2, graphs=True)
    _compare(e_stage, e_strip)


def test_strip_options_validated():
    from polarism_amd.engine import Rk4Engine
    e_stage, _, _ = _engines(SIZES[:1], "single", "double", 16)
    for bad in (32, 40, 20):
        with pytest.raises(ValueError, match="strip_h"):
            Rk4Engine(e_stage.table, e_stage.pump, "double", "single", whole_step=True, strip_h=bad)
    with pytest.raises(ValueError, match="whole-step"):
        Rk4Engine(e_stage.table, e_stage.pump, "double", "single", strip_h=16)


def _spinor_engines(sizes, reservoir, res_components, noise_strength=0.0):
    """Spinor members (psi+ and psi-, TE-TM, Zeeman, alpha_2, kinetic relaxation) on periodic grids
    in single precision: the stage launches and the spinor whole-step kernel."""
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.device import DeviceArray, init
    from polarism_amd.engine import Rk4Engine
    from polarism_amd.layout import MemberSpec, MemberTable
    init()
    p = PhysicsConstants()
    p.g_I = 0.0005
    p.kinetic_relaxation_eta = 0.02
    p.te_tm = 0.05
    p.zeeman = 0.02
    p.alpha_2 = -1e-4
    rng = np.random.default_rng(11)
    specs, envs, off = [], [], 0
    for i, (nx, ny) in enumerate(sizes):
        specs.append(MemberSpec(nx=nx, ny=ny, dx=0.5, dy=0.6, physics=p, reservoir=reservoir, components=2,
                                res_components=res_components, pump_off=off, seed=41 + i,
                                noise_strength=noise_strength, dt=0.05, circular_polarization=0.3))
        envs.append(rng.uniform(0.0, 1.0, nx * ny))
        off += nx * ny
    pump = DeviceArray.from_numpy(np.concatenate(envs).astype(np.float32))
    coef = DeviceArray.from_numpy(np.tile(rng.uniform(0.02, 0.08, len(sizes)), (4, 1)).astype(np.float32))
    e_stage = Rk4Engine(MemberTable(specs, 16), pump, "single", reservoir)
    e_step = Rk4Engine(MemberTable(specs, 16), pump, "single", reservoir, whole_step=True)
    total = e_stage.table.total_points
    psi = (0.3 * (rng.standard_normal(2 * total) + 1j * rng.standard_normal(2 * total))).astype(np.complex64)
    res = [rng.uniform(0.0, 3.0, res_components * total).astype(np.float32) for _ in range(e_stage.nres)]
    for e in (e_stage, e_step):
        e.set_dt(0.05)
        e.psi[0].upload(psi)
        for f in range(e.nres):
            e.res[f].upload(res[f])
    return e_stage, e_step, coef


@pytest.mark.parametrize("reservoir", ["none", "single", "double", "quadratic-double"])
@pytest.mark.parametrize("res_components", [2, 1])
def test_spinor_whole_step_equals_stage_launches_bitwise(reservoir, res_components):
    e_stage, e_step, coef = _spinor_engines(SIZES, reservoir, res_components)
    for e in (e_stage, e_step):
        e.run_steps(5, coef, True, False, 0, graphs=False)
    _compare(e_stage, e_step)


@pytest.mark.parametrize("res_components", [2, 1])
def test_spinor_whole_step_graph_chunks_noise_and_retired_members(res_components):
    """A graph-captured chunk, then single launches, with both Wigner streams fused in, on a tile
    map that covers a subset of the members (the others are left untouched in each field buffer,
    so the step count is even: the result is back in the buffer that holds their values)."""
    e_stage, e_step, coef = _spinor_engines(SIZES, "single", res_components, noise_strength=1.0)
    keep = [0, 1, 3]
    for e in (e_stage, e_step):
        e.set_active_tiles(e.table.tiles_of(np.asarray(keep)))
        e.run_steps(6, coef, True, True, 3, graphs=True)
        e.run_steps(2, coef, True, True, 9, graphs=False)
    _compare(e_stage, e_step)


def test_spinor_whole_step_scope():
    """Double precision and closed intervals keep the stage kernels; the engine says so."""
    from polarism_amd.engine import Rk4Engine
    e_stage, _, _ = _spinor_engines(SIZES[:1], "single", 2)
    with pytest.raises(ValueError, match="spinor"):
        Rk4Engine(e_stage.table, e_stage.pump, "double", "single", whole_step=True)
