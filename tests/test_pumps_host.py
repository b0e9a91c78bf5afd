"""Pump profiles: the oracle's lattice / ring / super-Gaussian restatements pinned by hand-checked
site lists, nearest-neighbour geometry, defects, rotation and the disorder draws; the package's
profiles equal the oracle's (profiles/builtin_profiles.py:28-240).  No GPU."""
import numpy as np
import pytest

from oracle import gpe_oracle as O
from polarism_amd import pump as P


def _grid(n=96, L=60.0):
    return O.Grid(n, n, L, L)


def test_square_lattice_sites_by_hand():
    """a = 10 on a 40 x 40 box: kept where |x|, |y| <= 20 + 10, i.e. x, y in {-30, ..., 30}."""
    sites = O.lattice_sites(40.0, 40.0, "square", 10.0)
    xs = sorted({round(x, 9) for x in sites[:, 3]})
    assert xs == [-30, -20, -10, 0, 10, 20, 30] and len(sites) == 49
    # order: i outer, then j
    assert list(sites[:3, 0]) == [-3, -3, -3] and list(sites[:3, 1]) == [-3, -2, -1]


@pytest.mark.parametrize("kind,nn,coord", [("square", 1.0, [4]), ("triangular", 1.0, [6]),
                                           ("honeycomb", 1 / np.sqrt(3), [3]), ("kagome", 1.0, [4]),
                                           ("lieb", 1.0, [4, 2])])
def test_lattice_nearest_neighbours(kind, nn, coord):
    """Nearest-neighbour distance (in units of the constant) and coordination number of sites
    well inside the box: the textbook values of each lattice."""
    a = 3.0
    s = O.lattice_sites(60.0, 60.0, kind, a)
    pts = s[:, 3:5]
    inner = np.flatnonzero(np.all(np.abs(pts) < 15.0, axis=1))
    d = np.sqrt(((pts[inner, None, :] - pts[None, :, :]) ** 2).sum(-1))
    d[d < 1e-9] = np.inf
    assert np.allclose(d.min(axis=1), nn * a)
    counts = sorted(set(int(c) for c in (np.abs(d - nn * a) < 1e-6).sum(axis=1)))
    assert counts == sorted(coord)


def test_defects_and_rotation():
    s0 = O.lattice_sites(40.0, 40.0, "honeycomb", 4.0)
    s1 = O.lattice_sites(40.0, 40.0, "honeycomb", 4.0, defects=[(0, 0, 1), (1, 0, 0)])
    assert len(s1) == len(s0) - 2
    assert not np.any(np.all(s1[:, :3] == [0, 0, 1], axis=1))
    # a square lattice turned by 90 degrees covers the same positions
    a = {tuple(np.round(r, 6)) for r in O.lattice_sites(30.0, 30.0, "square", 5.0)[:, 3:5]}
    b = {tuple(np.round(r, 6)) for r in O.lattice_sites(30.0, 30.0, "square", 5.0, rotation=90.0)[:, 3:5]}
    assert a == b


def test_disorder_draws_and_supergaussian_order_one():
    g = _grid(48, 40.0)
    flat = O.lattice_profile(g, "square", 8.0, 1.5)
    dis = O.lattice_profile(g, "square", 8.0, 1.5, disorder=0.2, seed=7)
    sites = O.lattice_sites(g.lx, g.ly, "square", 8.0)
    depth = 1.0 + np.random.default_rng(7).normal(0.0, 0.2, len(sites))
    manual = sum(dd * np.exp(-0.5 * ((g.X - x) ** 2 + (g.Y - y) ** 2) / 1.5 ** 2) for (_, _, _, x, y), dd in
                 zip(sites, depth))
    np.testing.assert_allclose(dis, manual, rtol=1e-12)
    assert not np.allclose(dis, flat)
    np.testing.assert_allclose(O.supergaussian_profile(g, 5.0, 1.0), O.gaussian_profile(g, 5.0), rtol=1e-15)
    np.testing.assert_allclose(O.lattice_profile(g, "square", 8.0, 1.5, well="supergaussian", order=1.0), flat,
                               rtol=1e-15)


@pytest.mark.parametrize("spec", [
    {"lattice": "square", "constant": 8.0, "width": 2.0},
    {"lattice": "triangular", "constant": 7.0, "width": 2.0, "rotation": 15.0},
    {"lattice": "honeycomb", "constant": 9.0, "width": 1.5, "disorder": 0.1, "seed": 3, "defects": [[0, 0, 0]]},
    {"lattice": "kagome", "constant": 6.0, "width": 1.2, "well": "supergaussian", "order": 2.0},
    {"lattice": "lieb", "constant": 5.0, "width": 1.0, "x0": 2.0, "y0": -1.0},
])
def test_package_lattice_equals_oracle(spec):
    g = _grid()
    mine = P.profile_lattice(g, dict(spec))
    kw = {k: v for k, v in spec.items() if k not in ("lattice", "constant", "width")}
    kw["defects"] = [tuple(d) for d in kw.get("defects", [])]
    ref = O.lattice_profile(g, spec["lattice"], spec["constant"], spec["width"], **kw)
    np.testing.assert_allclose(mine, ref, rtol=1e-12, atol=1e-14)


def test_package_ring_and_supergaussian_equal_oracle():
    g = _grid()
    np.testing.assert_allclose(P.profile_ring(g, {"radius": 15.0, "width": 3.0, "x0": 1.0}),
                               O.ring_profile(g, 15.0, 3.0, x0=1.0), rtol=1e-14)
    np.testing.assert_allclose(P.profile_supergaussian(g, {"sigma": 9.0, "order": 3.0}),
                               O.supergaussian_profile(g, 9.0, 3.0), rtol=1e-14)


@pytest.mark.parametrize("bad", [{"lattice": "hex", "constant": 1.0, "width": 1.0},
                                 {"lattice": "square", "constant": -1.0, "width": 1.0},
                                 {"lattice": "square", "constant": 1.0, "width": 1.0, "order": 2.0},
                                 {"lattice": "square", "constant": 1.0, "width": 1.0, "defects": [[0, 0, 1]]},
                                 {"lattice": "square", "constant": 1.0, "width": 1.0, "seed": -1}])
def test_lattice_validation(bad):
    with pytest.raises(ValueError):
        P.profile_lattice(_grid(16, 10.0), bad)


def test_training_pump_generators_follow_the_reference_profiles():
    """The surrogate's random pumps (polarism_amd.pumps.random_pump) are the reference profiles:
    a ring is builtin ring (r0, w); a lattice is the builtin square lattice of Gaussian wells
    (constant a, width w, origin (ox, oy)) times a Gaussian window; a mixture is a weighted sum
    of builtin Gaussians.  Each map is normalised to peak 1."""
    from polarism_amd.pumps import coords, random_pump
    n, L = 128, 128.0
    X, Y = coords(n, n, L, L)
    g = O.Grid(n, n, L, L)
    np.testing.assert_array_equal(X, g.X)
    for kind, seed in (("ring", 1), ("lattice", 2), ("lattice", 5), ("gaussians", 3)):
        m, desc = random_pump(np.random.default_rng(seed), n, n, L, L, kind)
        if kind == "ring":
            ref = O.ring_profile(g, desc["radius"], desc["width"])
        elif kind == "lattice":
            # the generator draws the origin right after period, width and window
            rng = np.random.default_rng(seed)
            a = rng.uniform(6, 15)
            width = a * rng.uniform(0.3, 0.45)
            window = rng.uniform(25, 40)
            ox, oy = rng.uniform(0, a, size=2)
            assert np.isclose(a, desc["period"]) and np.isclose(width, desc["width"])
            ref = O.lattice_profile(g, "square", a, width, x0=ox, y0=oy) * O.gaussian_profile(g, window)
        else:
            ref = sum(w * O.gaussian_profile(g, s, x0, y0) for s, x0, y0, w in desc["spots"])
        np.testing.assert_allclose(m, ref / ref.max(), rtol=1e-11, atol=1e-13)
