"""Result storage on the host: NPY / JSON layouts of format 3, dataset dims and units,
batching, static fields, accumulated values, the reader, and the formats this build
refuses (results/storage/*.py, docs/reference/results-format.md)."""
import json

import numpy as np
import pytest

from polarism_amd.accumulators import AccumulatedResult
from polarism_amd.config import Config
from polarism_amd.results import (JSONStorage, NPYStorage, ResultNode, StorageContext, create_run_directory,
                                  load_results, storage_names)
from polarism_amd.solver_api import SolverCompatibilityError


def _frames(storage, count, members=None, comps=1, seed=0):
    rng = np.random.default_rng(seed)
    nodes = [ResultNode("density", "|ψ|²"), ResultNode("norm", "N(t)", is_field=False),
             ResultNode("pump", "Pump", static=True)]
    lead = () if members is None else (members,)
    comp = () if comps == 1 else (comps,)
    pump = rng.random((4, 6))
    written = []
    for i in range(count):
        dens = rng.random((*lead, *comp, 4, 6)).astype(np.float32)
        norm = dens.reshape(*lead, -1).sum(axis=-1) if members else float(dens.sum())
        cached = {"density": dens, "norm": norm,
                  "pump": np.broadcast_to(pump, (*lead, 4, 6)).copy() if members else pump}
        groups = {"P_lasers": {"L0": np.full(members, 0.5 * i) if members else 0.5 * i}}
        storage.add_to_batch(0.25 * i, nodes, cached=cached, scalar_groups=groups)
        written.append((0.25 * i, dens, norm))
    return written, pump


@pytest.mark.parametrize("cls", [NPYStorage, JSONStorage])
def test_round_trip_through_the_reader(tmp_path, cls):
    cfg = Config()
    x = np.linspace(-1, 1, 6)
    y = np.linspace(-2, 2, 4)
    st = cls(tmp_path, 3, cfg, x, y, StorageContext())
    written, pump = _frames(st, 7)
    acc = {"g2": AccumulatedResult("g2", np.ones((4, 6)), ("y", "x"), "", "g2", True, False)}
    st.write_accumulated(acc, {"g2": {"samples": 5, "t_first": 0.0, "t_last": 1.0}})
    st.set_stop_times([1.5])
    st.finalize()
    kind = "npz" if cls is NPYStorage else "json"
    # 7 frames in batches of 3 -> 3 batch files
    assert len(list(tmp_path.glob(f"batch_*.{kind}"))) == 3
    r = load_results(tmp_path)
    np.testing.assert_array_equal(r.time, [w[0] for w in written])
    np.testing.assert_allclose(r.fields["density"], np.stack([w[1] for w in written]), rtol=1e-7)
    np.testing.assert_allclose(r.scalars["norm"], [w[2] for w in written], rtol=1e-12)
    np.testing.assert_allclose(r.scalar_groups["P_lasers"]["L0"], 0.5 * np.arange(7))
    np.testing.assert_allclose(r.static["pump"], pump)
    np.testing.assert_array_equal(r.accumulated["g2"], np.ones((4, 6)))
    assert r.accumulated_meta["g2"]["samples"] == 5 and r.accumulated_meta["g2"]["linear"] is False
    assert r.attrs["format_version"] == 3 and r.attrs["stop_time"] == [1.5]
    assert r.attrs["datasets"]["fields/density"] == {"label": "|ψ|²", "dims": ["time", "y", "x"],
                                                     "units": "µm^-2"}
    assert r.attrs["datasets"]["scalars/norm"]["dims"] == ["time"]
    assert r.attrs["datasets"]["static/pump"]["dims"] == ["y", "x"]
    assert r.config["grid"]["nx"] == cfg.grid.nx
    if cls is NPYStorage:
        np.testing.assert_array_equal(r.x, x)
        np.testing.assert_allclose(r.kx, np.fft.fftshift(2 * np.pi * np.fft.fftfreq(6, d=x[1] - x[0])))


def test_batched_spinor_dims_members_and_static_sharing(tmp_path):
    cfg = Config()
    cfg.physics.spin = "spinor"
    ctx = StorageContext(members=3, member_indices=(0, 1, 2), member_seeds=(7, 8, 9))
    st = NPYStorage(tmp_path, 100, cfg, np.arange(6.0), np.arange(4.0), ctx)
    _frames(st, 2, members=3, comps=2)
    st.finalize()
    meta = json.loads((tmp_path / "metadata.json").read_text())
    assert meta["members"] == 3 and meta["member_seeds"] == [7, 8, 9]
    assert meta["datasets"]["fields/density"]["dims"] == ["time", "member", "component", "y", "x"]
    assert meta["datasets"]["scalars/norm"]["dims"] == ["time", "member"]
    # a pump shared by every member is stored once without the member axis
    assert meta["datasets"]["static/pump"]["dims"] == ["y", "x"]
    assert np.load(tmp_path / "static_pump.npy").shape == (4, 6)
    np.testing.assert_array_equal(np.load(tmp_path / "member_index.npy"), [0, 1, 2])
    with np.load(tmp_path / "batch_000000.npz") as z:
        assert z["field_density"].shape == (2, 3, 2, 4, 6)
        assert z["scalar_norm"].shape == (2, 3)


def test_dataset_metadata_rules():
    cfg = Config()
    cfg.physics.spin = "spinor"
    cfg.reservoir.components = 2
    st = NPYStorage.__new__(NPYStorage)
    st.ndim, st.components, st.reservoir_components, st.members, st.node_metadata = 2, 2, 2, None, {}
    assert st.dataset_metadata("stokes", field=True)["dims"] == ["time", "stokes", "y", "x"]
    assert st.dataset_metadata("nA", field=True)["dims"] == ["time", "component", "y", "x"]
    assert st.dataset_metadata("nA", field=True)["units"] == "µm^-2"
    assert st.dataset_metadata("psi", field=True)["units"] == "µm^-1"
    assert st.dataset_metadata("momentum_density", field=True)["dims"] == ["time", "component", "ky", "kx"]
    assert st.dataset_metadata("density_plus", field=True)["dims"] == ["time", "y", "x"]
    st.reservoir_components = 1
    assert st.dataset_metadata("N_R", field=True)["dims"] == ["time", "y", "x"]


def test_batches_split_by_bytes(tmp_path, monkeypatch):
    import polarism_amd.results as R
    monkeypatch.setattr(R, "MAX_BATCH_BYTES", 200)
    st = NPYStorage(tmp_path, 1000, Config(), None, None, StorageContext())
    _frames(st, 5)   # 96 bytes of density per frame: the third frame passes 200 bytes
    st.finalize()
    assert len(list(tmp_path.glob("batch_*.npz"))) == 2
    with np.load(tmp_path / "batch_000000.npz") as z:
        assert z["time"].shape == (3,)


def test_complex_fields_in_json(tmp_path):
    st = JSONStorage(tmp_path, 10, Config(), None, None, StorageContext())
    psi = (np.arange(24).reshape(4, 6) + 1j * np.arange(24).reshape(4, 6)[::-1]).astype(np.complex64)
    st.add_to_batch(0.0, [ResultNode("psi")], cached={"psi": psi}, scalar_groups={})
    st.finalize()
    r = load_results(tmp_path)
    np.testing.assert_array_equal(r.fields["psi"][0], psi)


def test_format_selection_and_refusals(tmp_path):
    cfg = Config()
    cfg.result.formats = ("json",)
    cfg.result.save_npy = True
    assert storage_names(cfg.result) == ["json", "npy"]
    cfg.result.save_hdf5 = True
    with pytest.raises(SolverCompatibilityError, match="h5py"):
        storage_names(cfg.result)
    cfg.result.save_hdf5 = False
    cfg.result.formats = ("parquet",)
    with pytest.raises(ValueError, match="Unknown result storage"):
        storage_names(cfg.result)
    (tmp_path / "results.h5").write_bytes(b"")
    with pytest.raises(SolverCompatibilityError):
        load_results(tmp_path)


def test_run_directories_are_unique(tmp_path):
    a = create_run_directory(tmp_path)
    b = create_run_directory(tmp_path)
    assert a != b and a.is_dir() and b.is_dir()
