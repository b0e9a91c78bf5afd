"""Batched 2-D LDS Stockham FFT against numpy.fft (FP64 to ~1e-15, FP32 to ~1e-7)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(16, 16), (64, 96), (96, 128), (192, 64), (256, 256), (384, 48), (512, 512), (768, 32), (1024, 1024)]


def _plan(nx, ny, prec):
    from polarism_amd.solvers.split_step_hip import FFTPlan
    return FFTPlan(nx, ny, 1.0, 1.0, prec)


@pytest.mark.parametrize("ny,nx", SIZES)
@pytest.mark.parametrize("precision,tol", [("double", 1e-14), ("single", 2e-6)])
def test_fft2_matches_numpy(ny, nx, precision, tol):
    from polarism_amd import _native as N
    from polarism_amd.device import DeviceArray, init
    init()
    rng = np.random.default_rng(ny * 1000 + nx)
    B = 3
    cdt = np.complex64 if precision == "single" else np.complex128
    x = (rng.standard_normal((B, ny, nx)) + 1j * rng.standard_normal((B, ny, nx))).astype(cdt)
    plan = _plan(nx, ny, precision)
    d_in = DeviceArray.from_numpy(x)
    d_out = DeviceArray.empty(x.shape, cdt)
    N.call("pgpe_fft2", plan.handle, d_in.ptr, d_out.ptr, B, 0, None)
    ref = np.fft.fft2(x.astype(np.complex128), axes=(-2, -1))
    got = d_out.get()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < tol, err
    # inverse, in place
    N.call("pgpe_fft2", plan.handle, d_out.ptr, d_out.ptr, B, 1, None)
    back = d_out.get()
    err_b = np.linalg.norm(back - x) / np.linalg.norm(x)
    assert err_b < tol, err_b


def test_fft2_plane_wave_and_parseval():
    from polarism_amd import _native as N
    from polarism_amd.device import DeviceArray, init
    init()
    ny, nx = 96, 256
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    f = np.exp(2j * np.pi * (5 * x / nx + 7 * y / ny))
    plan = _plan(nx, ny, "double")
    d = DeviceArray.from_numpy(f.reshape(1, ny, nx))
    N.call("pgpe_fft2", plan.handle, d.ptr, d.ptr, 1, 0, None)
    got = d.get()[0]
    expected = np.zeros_like(f)
    expected[7, 5] = nx * ny
    assert np.max(np.abs(got - expected)) < 1e-9
    assert abs(np.sum(np.abs(got) ** 2) / (nx * ny) - np.sum(np.abs(f) ** 2)) < 1e-8 * nx * ny


def test_plan_rejects_unsupported_size():
    from polarism_amd.device import init
    init()
    with pytest.raises(RuntimeError, match="sizes"):
        _plan(100, 64, "double")
    # 6-point complex128 rows do not fit a row group's LDS: rejected by the plan, not at launch
    with pytest.raises(RuntimeError, match="short"):
        _plan(6, 64, "double")
    _plan(6, 64, "single")


@pytest.mark.parametrize("ny,nx,B", [(2048, 16, 40), (2048, 64, 64), (2048, 12, 400), (4096, 32, 64),
                                     (1024, 32, 96), (1024, 3, 700)])
@pytest.mark.parametrize("precision,tol", [("double", 1e-14), ("single", 2e-6)])
def test_fft2_tall_columns_in_large_batches(ny, nx, B, precision, tol):
    """Large batches take wide column groups (up to 64 elements per thread, 64-byte row segments).
    Batches of >= 512 groups of complex64 columns of 2048 / 4096 points or complex128 columns of
    1024 points split each column in two halves (fft_cols_half_kernel), including groups that hold
    fewer columns than the full group (nx = 12, 3)."""
    from polarism_amd import _native as N
    from polarism_amd.device import DeviceArray, init
    init()
    rng = np.random.default_rng(ny + nx + B)
    cdt = np.complex64 if precision == "single" else np.complex128
    x = (rng.standard_normal((B, ny, nx)) + 1j * rng.standard_normal((B, ny, nx))).astype(cdt)
    plan = _plan(nx, ny, precision)
    d_in = DeviceArray.from_numpy(x)
    d_out = DeviceArray.empty(x.shape, cdt)
    N.call("pgpe_fft2", plan.handle, d_in.ptr, d_out.ptr, B, 0, None)
    ref = np.fft.fft2(x.astype(np.complex128), axes=(-2, -1))
    err = np.linalg.norm(d_out.get() - ref) / np.linalg.norm(ref)
    assert err < tol, err


@pytest.mark.parametrize("ny,nx,B", [(512, 512, 1), (64, 256, 1), (2048, 64, 1), (2048, 64, 64)])
def test_fp32_round_trips_keep_the_norm(ny, nx, B):
    """120 forward + inverse FP32 round trips: the energy drifts by less than 1e-6 and the field
    by less than 1e-5 (rounding only).  FP32 twiddles rounded without tails lose ~8e-8 of the
    energy per 2-D transform on 512^2 (1.5e-5 over these round trips); numpy's complex64
    transform loses 7e-6."""
    from polarism_amd import _native as N
    from polarism_amd.device import DeviceArray, init
    init()
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((B, ny, nx)) + 1j * rng.standard_normal((B, ny, nx))).astype(np.complex64)
    plan = _plan(nx, ny, "single")
    d = DeviceArray.from_numpy(x)
    for _ in range(120):
        N.call("pgpe_fft2", plan.handle, d.ptr, d.ptr, B, 0, None)
        N.call("pgpe_fft2", plan.handle, d.ptr, d.ptr, B, 1, None)
    got = d.get().astype(np.complex128)
    x64 = x.astype(np.complex128)
    drift = np.sum(np.abs(got) ** 2) / np.sum(np.abs(x64) ** 2) - 1
    err = np.linalg.norm(got - x64) / np.linalg.norm(x64)
    print(f"{B} x {ny}x{nx}: 120 FP32 round trips: energy drift {drift:.2e}, field error {err:.2e}")
    assert abs(drift) < 1e-6 and err < 1e-5
