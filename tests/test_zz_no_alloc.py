"""pgpe_density_delta_noalloc (partials in the library's built-in scratch) against
pgpe_density_delta (partials in a caller workspace), through the C ABI: two calls in a row, the
second on a smaller table, scalar and spinor, single and double; plus the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(sizes, components, precision, seed):
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.device import DeviceArray
    from polarism_amd.layout import MemberSpec, MemberTable
    rng = np.random.default_rng(seed)
    c = np.complex64 if precision == "single" else np.complex128
    r = np.float32 if precision == "single" else np.float64
    tab = MemberTable([MemberSpec(nx, ny, 0.5, 0.5, PhysicsConstants(), components=components,
                                  res_components=2 if components == 2 else 1) for nx, ny in sizes], 16)
    n = tab.total_points
    psi = (rng.standard_normal(n * components) + 1j * rng.standard_normal(n * components)).astype(c)
    old = np.abs(rng.standard_normal(n)).astype(r)
    return tab, DeviceArray.from_numpy(psi), old


def _run_both(tab, psi, old_host, components, precision, tile_map=None, num_tiles=None):
    from polarism_amd import _native as N
    from polarism_amd.device import DeviceArray
    prec = 0 if precision == "single" else 1
    tmap = tab.tile_map_ptr if tile_map is None else tile_map.ptr
    nt = tab.num_tiles if num_tiles is None else num_tiles
    old_a = DeviceArray.from_numpy(old_host)
    old_b = DeviceArray.from_numpy(old_host)
    partials = DeviceArray.zeros((3 * tab.num_tiles,), np.float64)
    sums_a = DeviceArray.zeros((3 * tab.members,), np.float64)
    sums_b = DeviceArray.zeros((3 * tab.members,), np.float64)
    N.call("pgpe_density_delta", psi.ptr, old_a.ptr, prec, components, tab.table_ptr, tmap, nt, tab.members,
           partials.ptr, sums_a.ptr, 1, None)
    lib = N.lib()
    rc = lib.pgpe_density_delta_noalloc(psi.ptr, old_b.ptr, prec, components, tab.table_ptr, tmap, nt, tab.members,
                                        tab.num_tiles, sums_b.ptr, 1, None)
    assert rc == 0, N.last_error()
    return sums_a.get().reshape(-1, 3), sums_b.get().reshape(-1, 3), old_a.get(), old_b.get()


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("components", [1, 2])
def test_noalloc_equals_workspace_variant_on_two_calls(precision, components):
    from polarism_amd.device import init
    init()
    print(f"no-alloc density delta: precision={precision} components={components}", flush=True)
    big = [(256, 256), (100, 37), (512, 96), (33, 130)]
    small = [(64, 48), (40, 40)]
    for sizes, seed in ((big, 1), (small, 2)):
        tab, psi, old = _case(sizes, components, precision, seed)
        a, b, oa, ob = _run_both(tab, psi, old, components, precision)
        np.testing.assert_array_equal(b, a)
        np.testing.assert_array_equal(ob, oa)


def test_noalloc_retired_members_sum_to_zero():
    """A partial tile map: running members agree with the workspace variant; members with no
    tiles in the map read the cleared scratch, not an earlier call's partials."""
    from polarism_amd.device import DeviceArray, init
    init()
    tab, psi, old = _case([(128, 128), (96, 64), (64, 64)], 1, "double", 3)
    _run_both(tab, psi, old, 1, "double")                     # fills the scratch with this table's partials
    keep = tab.tiles_of([0, 2])
    tmap = DeviceArray.from_numpy(keep)
    a, b, _, _ = _run_both(tab, psi, old, 1, "double", tile_map=tmap, num_tiles=int(keep.shape[0]))
    np.testing.assert_array_equal(b[[0, 2]], a[[0, 2]])
    np.testing.assert_array_equal(b[1], 0.0)


def test_noalloc_rejects_tables_beyond_the_scratch():
    from polarism_amd import _native as N
    from polarism_amd.device import init
    init()
    tab, psi, old = _case([(64, 64)], 1, "single", 4)
    from polarism_amd.device import DeviceArray
    o = DeviceArray.from_numpy(old)
    sums = DeviceArray.zeros((3,), np.float64)
    rc = N.lib().pgpe_density_delta_noalloc(psi.ptr, o.ptr, 0, 1, tab.table_ptr, tab.tile_map_ptr, tab.num_tiles, 1,
                                            65537, sums.ptr, 1, None)
    assert rc != 0 and "65536" in N.last_error()
