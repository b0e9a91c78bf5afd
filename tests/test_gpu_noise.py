"""GPU Philox words, Box-Muller normals and Wigner increments against the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import gpe_oracle as O

pytestmark = pytest.mark.gpu


def _dev():
    from polarism_amd import device
    device.init()
    return device


def test_philox_words_bit_exact():
    device = _dev()
    from polarism_amd import _native as N
    counters = O.counter_words(np.arange(4096), (1 << 40) + 7)
    key = 0x123456789ABCDEF0
    expected = O.philox4x32(counters, key)
    src = device.DeviceArray.from_numpy(counters)
    out = device.DeviceArray.empty(counters.shape, np.uint32)
    N.call("pgpe_philox_words", src.ptr, out.ptr, key, counters.shape[0], None)
    np.testing.assert_array_equal(out.get(), expected)


@pytest.mark.parametrize("key,counter,expected", [
    (0, (0, 0, 0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    (0xFFFFFFFFFFFFFFFF, (0xFFFFFFFF,) * 4, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    (0x299F31D0A4093822, (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_random123_vectors_gpu(key, counter, expected):
    device = _dev()
    from polarism_amd import _native as N
    src = device.DeviceArray.from_numpy(np.asarray([counter], dtype=np.uint32))
    out = device.DeviceArray.empty((1, 4), np.uint32)
    N.call("pgpe_philox_words", src.ptr, out.ptr, key, 1, None)
    assert tuple(int(v) for v in out.get()[0]) == expected


@pytest.mark.parametrize("precision,tol", [("single", 2e-6), ("double", 2e-14)])
def test_normal_pairs_match_oracle(precision, tol):
    device = _dev()
    from polarism_amd import _native as N
    n = 65536
    key, step = 0xDEADBEEF12345, 123456789
    cdt = np.complex64 if precision == "single" else np.complex128
    out = device.DeviceArray.empty((n,), cdt)
    N.call("pgpe_normal_pairs", out.ptr, 0 if precision == "single" else 1, key, step, 1, n, None)
    re, im = O.normal_pair(O.philox4x32(O.counter_words(np.arange(n), step, 1), key), precision)
    got = out.get()
    np.testing.assert_allclose(got.real, re, rtol=tol, atol=tol)
    np.testing.assert_allclose(got.imag, im, rtol=tol, atol=tol)


@pytest.mark.parametrize("precision,tol", [("single", 2e-6), ("double", 2e-14)])
def test_wigner_add_matches_oracle(precision, tol):
    from polarism_amd.layout import MemberSpec, MemberTable
    from polarism_amd import _native as N
    device = _dev()
    real = np.float32 if precision == "single" else np.float64
    cdt = np.complex64 if precision == "single" else np.complex128
    ny, nx = 40, 70           # ragged: not a multiple of the 64 x 16 tile
    p = O.Physics()
    B = 3
    seeds = [11, 22, 2 ** 40 + 5]
    strength = [1.0, 0.0, 2.5]
    dt, dx, dy = 0.01, 0.5, 0.25
    rng = np.random.default_rng(1)
    psi = (rng.standard_normal((B, 1, ny, nx)) + 1j * rng.standard_normal((B, 1, ny, nx))).astype(cdt) * 1e-3
    active = rng.random((B, 1, ny, nx)).astype(real)
    specs = [MemberSpec(nx, ny, dx, dy, p, seed=seeds[m], noise_strength=strength[m], dt=dt) for m in range(B)]
    tab = MemberTable(specs, 16)
    d_psi = device.DeviceArray.from_numpy(psi)
    d_act = device.DeviceArray.from_numpy(active)
    N.call("pgpe_wigner_add", d_psi.ptr, d_act.ptr, 1, tab.table_ptr, tab.tile_map_ptr, tab.num_tiles,
           B, 0 if precision == "single" else 1, 77, None)
    ref = psi.copy()
    O.wigner_add(ref, active, seeds, [p.gamma_C] * B, [p.R] * B, strength, dt, dx * dy, 77, precision)
    np.testing.assert_allclose(d_psi.get(), ref, rtol=tol, atol=tol * 1e-3)
