"""Autotuner cache logic (host only: the timer is injected)."""
from polarism_amd.autotune import Autotuner, LaunchChoice, candidates, tune_key


def test_key_depends_on_size_multiset_not_order():
    a = tune_key("rk4-fdm-hip", "dev", "single", "double", 1, [(64, 64), (128, 96), (64, 64)])
    b = tune_key("rk4-fdm-hip", "dev", "single", "double", 1, [(128, 96), (64, 64), (64, 64)])
    c = tune_key("rk4-fdm-hip", "dev", "single", "double", 1, [(128, 96), (64, 64)])
    d = tune_key("rk4-fdm-hip", "dev", "double", "double", 1, [(128, 96), (64, 64), (64, 64)])
    assert a == b and a != c and a != d


def test_candidates_respect_kernel_limits():
    assert LaunchChoice("step", 64) in candidates("single", 1)
    assert LaunchChoice("step", 64) not in candidates("double", 1)
    assert all(c.form == "stage" for c in candidates("single", 2))
    assert LaunchChoice("strip", 32) in candidates("single", 1)
    assert LaunchChoice("strip", 32) not in candidates("double", 1)
    assert LaunchChoice("strip", 40) in candidates("single", 1)
    assert LaunchChoice("strip", 40) not in candidates("double", 1)
    assert LaunchChoice("strip", 24) in candidates("double", 1)
    assert LaunchChoice("strip", 8) in candidates("double", 1) and LaunchChoice("strip", 8) in candidates("single", 1)


def test_choice_geometry():
    """Row strips run the whole-step kernel on a member table of 16-row tiles (the reductions'
    tiles) with their own strip height; the other forms use their height for the table."""
    s = LaunchChoice("strip", 24)
    assert s.whole_step and s.table_h == 16 and s.strip_h == 24
    t = LaunchChoice("step", 32)
    assert t.whole_step and t.table_h == 32 and t.strip_h is None
    g = LaunchChoice("stage", 8)
    assert not g.whole_step and g.table_h == 8 and g.strip_h is None


def test_choice_is_fastest_and_persisted(tmp_path):
    path = tmp_path / "tune.json"
    times = {LaunchChoice("stage", 8): 3.0, LaunchChoice("stage", 16): 2.0, LaunchChoice("step", 32): 1.0}
    calls = []

    def timer(opt):
        calls.append(opt)
        return times[opt]

    t = Autotuner(path)
    best = t.choose("k", list(times), timer)
    assert best == LaunchChoice("step", 32) and len(calls) == 3
    again = Autotuner(path).choose("k", list(times), lambda o: 1 / 0)     # cached: no timing
    assert again == best
    assert Autotuner(path).lookup("other") is None


def test_corrupt_cache_is_ignored(tmp_path):
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    t = Autotuner(path)
    assert t.lookup("k") is None
    assert t.choose("k", [LaunchChoice("stage", 16)], lambda o: 1.0) == LaunchChoice("stage", 16)


def test_sample_members_takes_the_whole_table_within_the_budget():
    from types import SimpleNamespace as S
    from polarism_amd.autotune import sample_members
    specs = [S(nx=256, ny=256) for _ in range(256)]            # 16.8 M points
    assert sample_members(specs, 1 << 25) == specs
    mixed = [S(nx=n, ny=n) for n in (128, 256, 384, 512) * 64]  # 31.5 M points
    sub = sample_members(mixed, 1 << 24)
    assert sum(s.nx * s.ny for s in sub) <= 1 << 24
    assert len(sub) > len(mixed) // 3                            # close to the budget, not a token sample
    assert {s.nx for s in sub} == {128, 256, 384, 512}           # the size mix is kept
    assert len(sample_members([S(nx=8192, ny=8192)], 1 << 20)) == 1
