"""Pins of the oracle by closed forms, invariants and brute force (CPU only).

Each test checks an oracle function against something the mathematics fixes,
chosen so that a dropped term, a wrong sign or a transposed index fails it.
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import gpe_oracle as O


def _zero_pump(grid, dtype=np.float64):
    z = np.zeros((1, 1) + grid.shape, dtype=dtype)
    return lambda t: z


# ---------------------------------------------------------------- Laplacian
def test_laplacian_cosine_eigenfunction_and_second_order():
    """The periodic 5-point stencil maps cos(kx x) cos(ky y) to the discrete symbol times itself."""
    errs = []
    for n in (32, 64):
        g = O.Grid(n, n // 2 * 3 // 2, 40.0, 30.0)
        kx, ky = 2 * np.pi * 3 / g.lx, 2 * np.pi * 2 / g.ly
        f = np.cos(kx * g.X) * np.cos(ky * g.Y)
        lap = O.laplacian(f, g.dx, g.dy)
        sym = -(2 - 2 * np.cos(kx * g.dx)) / g.dx ** 2 - (2 - 2 * np.cos(ky * g.dy)) / g.dy ** 2
        np.testing.assert_allclose(lap, sym * f, atol=1e-12 * abs(sym))
        errs.append(abs(sym + kx ** 2 + ky ** 2))
    assert 3.8 < errs[0] / errs[1] < 4.2


def test_te_tm_stencil_reduces_to_dispersion_coupling():
    """-beta[(dxx - dyy) - 2i dxy] on a plane-wave psi_- equals beta(kx - i ky)^2 psi_- up to the stencil symbols."""
    g = O.Grid(64, 48, 64.0, 48.0)
    p = O.Physics(te_tm=0.05)
    kx, ky = 2 * np.pi * 2 / g.lx, 2 * np.pi * 3 / g.ly
    wave = np.exp(1j * (kx * g.X + ky * g.Y))
    sx = -(2 - 2 * np.cos(kx * g.dx)) / g.dx ** 2
    sy = -(2 - 2 * np.cos(ky * g.dy)) / g.dy ** 2
    sxy = -np.sin(kx * g.dx) * np.sin(ky * g.dy) / (g.dx * g.dy)
    for src in (0, 1):
        field = np.zeros((1, 2) + g.shape, dtype=complex)
        field[0, src] = wave
        out = O.te_tm_energy(p, field, g.dx, g.dy)
        dst = 1 - src
        sign = -1 if dst == 0 else 1          # psi_+ row gets -2i dxy, psi_- row +2i dxy
        expect = -p.te_tm * ((sx - sy) + sign * 2j * sxy) * wave
        np.testing.assert_allclose(out[0, dst], expect, atol=1e-12)
        np.testing.assert_allclose(out[0, src], 0, atol=1e-15)
    # continuum limit of the psi_+ row: beta (kx - i ky)^2, the off-diagonal of the dispersion matrix
    H = O.spinor_hamiltonian(g, p)
    i, j = 3, 2                                 # grid index of (ky, kx) = (3, 2) modes
    assert np.isclose(g.KX[i, j], kx) and np.isclose(g.KY[i, j], ky)
    cont = -p.te_tm * ((sx - sy) - 2j * sxy)
    # stencil symbols differ from kx, ky at O((k dx)^2); a wrong sign of the 2i dxy term is O(1) off
    assert abs(cont - H[0, 1][i, j]) < 5e-2 * abs(H[0, 1][i, j])


def test_te_tm_closed_interval_on_neumann_cosine_modes():
    """On the closed interval the stencil sees the field mirrored about the wall nodes.

    f = cos(a x_i) cos(b y_j) with a = pi m / (nx - 1), b = pi l / (ny - 1) is its own mirror
    extension (f[-1] = f[1], f[n] = f[n - 2]), so at every point, walls and corners included,
    dxx = (2 cos a - 2) / dx^2 f, dyy = (2 cos b - 2) / dy^2 f and
    dxy = sin a sin b sin(a i) sin(b j) / (dx dy).  A periodic wrap, an edge-repeating mirror,
    swapped axes or a flipped 2i dxy sign each miss these values at the walls or everywhere.
    """
    g = O.Grid(13, 9, 24.0, 12.0, boundary="closed-interval")
    p = O.Physics(te_tm=0.05)
    i = np.arange(g.nx)[None, :]
    j = np.arange(g.ny)[:, None]
    a, b = np.pi * 3 / (g.nx - 1), np.pi * 2 / (g.ny - 1)
    mode = np.cos(a * i) * np.cos(b * j)
    dxx = (2 * np.cos(a) - 2) / g.dx ** 2 * mode
    dyy = (2 * np.cos(b) - 2) / g.dy ** 2 * mode
    dxy = np.sin(a) * np.sin(b) * np.sin(a * i) * np.sin(b * j) / (g.dx * g.dy)
    for src in (0, 1):
        field = np.zeros((1, 2) + g.shape, dtype=complex)
        field[0, src] = (0.6 - 0.8j) * mode
        out = O.te_tm_energy(p, field, g.dx, g.dy, boundary="closed-interval")
        dst = 1 - src
        sign = -1 if dst == 0 else 1          # psi_+ row gets -2i dxy, psi_- row +2i dxy
        expect = -p.te_tm * ((dxx - dyy) + sign * 2j * dxy) * (0.6 - 0.8j)
        np.testing.assert_allclose(out[0, dst], expect, atol=1e-12)
        np.testing.assert_allclose(out[0, src], 0, atol=1e-15)
    # the periodic stencil on the same samples differs at the walls
    field = np.zeros((1, 2) + g.shape, dtype=complex)
    field[0, 1] = mode
    per = O.te_tm_energy(p, field, g.dx, g.dy)
    assert np.abs(per[0, 0] - O.te_tm_energy(p, field, g.dx, g.dy, boundary="closed-interval")[0, 0]).max() > 1e-3


# ---------------------------------------------------------------- spinor dispersion
@pytest.mark.parametrize("zeeman,te_tm", [(0.02, 0.05), (0.0, 0.05), (0.02, 0.0), (0.0, 0.0)])
def test_spinor_propagator_equals_matrix_exponential(zeeman, te_tm):
    g = O.Grid(8, 6, 10.0, 7.0)
    p = O.Physics(zeeman=zeeman, te_tm=te_tm)
    tau = 0.37
    M = O.spinor_propagator(g, p, tau)
    H = O.spinor_hamiltonian(g, p)
    for i in range(g.ny):
        for j in range(g.nx):
            ref = expm(-1j * H[:, :, i, j] * tau / p.hbar)
            np.testing.assert_allclose(M[:, :, i, j], ref, atol=1e-13)
    # Hermitian H: the propagator is unitary
    MM = np.einsum("abij,cbij->acij", M, M.conj())
    np.testing.assert_allclose(MM, np.eye(2)[:, :, None, None] * np.ones(g.shape), atol=1e-13)


def test_zeeman_only_spinor_rhs_equals_scalar_with_offset_potential():
    """+Z/2 on psi_+, -Z/2 on psi_-: the same sign as the dispersion matrix diagonal."""
    g = O.Grid(16, 16, 20.0, 20.0)
    p = O.Physics(zeeman=0.03)
    rng = np.random.default_rng(3)
    psi = (rng.standard_normal((1, 2) + g.shape) + 1j * rng.standard_normal((1, 2) + g.shape)) * 0.3
    res = (np.abs(rng.standard_normal((1, 2) + g.shape)),)
    pump = np.full((1, 1) + g.shape, 0.02)
    d, _ = O.fdm_rhs("single", p, g, psi, res, pump, pump_weights=np.ones((1, 2, 1, 1)), reservoir_components=2)
    for c, sgn in ((0, 1), (1, -1)):
        ds, _ = O.fdm_rhs("single", p.with_(zeeman=0.0), g, psi[:, c:c + 1], (res[0][:, c:c + 1],), pump,
                          potential=sgn * p.zeeman / 2)
        np.testing.assert_allclose(d[:, c], ds[:, 0], rtol=1e-13, atol=1e-16)


# ---------------------------------------------------------------- uniform closed forms
def _uniform_exact(p, t, amp=1.0):
    gamma = p.gamma_C
    phase = -p.g_C * amp ** 2 / (p.hbar * gamma) * (1 - np.exp(-gamma * t))
    return amp * np.exp(-gamma * t / 2 + 1j * phase)


def test_rk4_fdm_uniform_decay_is_fourth_order():
    g = O.Grid(4, 4, 4.0, 4.0)
    p = O.Physics(g_C=0.5, gamma_C=0.4)
    T = 2.0
    errs = []
    for dt in (0.2, 0.1):
        psi = np.ones((1, 1) + g.shape, dtype=complex)
        for s in range(int(round(T / dt))):
            psi, _ = O.rk4_fdm_step("none", p, g, psi, (), _zero_pump(g), s * dt, dt)
        errs.append(np.max(np.abs(psi - _uniform_exact(p, T))))
    assert errs[1] < 1e-5
    assert 13 < errs[0] / errs[1] < 19


@pytest.mark.parametrize("stepper", [O.split_step_cpu, O.split_step_kouter])
def test_split_step_exact_for_uniform_field(stepper):
    """The exact local step integrates uniform gain/loss and the self phase exactly."""
    g = O.Grid(8, 8, 8.0, 8.0)
    p = O.Physics(g_C=0.5, gamma_C=0.4)
    dt, n = 0.05, 40
    psi = np.full((1, 1) + g.shape, 0.7 + 0j)
    for s in range(n):
        psi, _ = stepper("none", p, g, psi, (), _zero_pump(g), s * dt, dt, s)
    np.testing.assert_allclose(psi, _uniform_exact(p, n * dt, 0.7), rtol=1e-12)


@pytest.mark.parametrize("stepper", [O.split_step_cpu, O.split_step_kouter])
def test_split_step_plane_wave_phase(stepper):
    g = O.Grid(16, 8, 20.0, 10.0)
    p = O.Physics(g_C=0.0, gamma_C=0.0)
    kx, ky = 2 * np.pi * 3 / g.lx, -2 * np.pi * 1 / g.ly
    psi = np.exp(1j * (kx * g.X + ky * g.Y))[None, None]
    psi0 = psi.copy()
    dt, n = 0.1, 25
    for s in range(n):
        psi, _ = stepper("none", p, g, psi, (), _zero_pump(g), s * dt, dt, s)
    omega = p.hbar * (kx ** 2 + ky ** 2) / (2 * p.m_eff)
    np.testing.assert_allclose(psi, psi0 * np.exp(-1j * omega * n * dt), atol=1e-12)


def test_split_step_orderings_differ_at_second_order():
    g = O.Grid(32, 32, 30.0, 30.0)
    p = O.Physics(g_C=0.05)
    env = O.gaussian_profile(g, 6.0)[None, None]
    pump = lambda t: 0.5 * env
    psi0 = np.full((1, 1) + g.shape, 0.5 + 0j) * (1 + 0.2 * np.cos(2 * np.pi * g.X / g.lx))
    diffs = []
    T = 1.0
    for dt in (0.02, 0.01):
        a = b = psi0
        ra = rb = (np.zeros((1, 1) + g.shape),)
        for s in range(int(round(T / dt))):
            a, ra = O.split_step_cpu("single", p, g, a, ra, pump, s * dt, dt, s)
            b, rb = O.split_step_kouter("single", p, g, b, rb, pump, s * dt, dt, s)
        diffs.append(O.rel_l2(ra[0], rb[0]))
    assert diffs[1] > 1e-11
    assert 3.5 < diffs[0] / diffs[1] < 4.5


# ---------------------------------------------------------------- reservoirs and thresholds
def _fixed_point_at_zero_density(kind, p, P):
    """Reservoir fixed point with psi = 0 by solving the linear system dn/dt = 0."""
    if kind == "single":
        return (P / p.gamma_R,)
    # double: [[-(gA + R_AI), R_IA], [R_AI, -(gI + R_IA)]] (nA, nI) = (0, -P)
    A = np.array([[-(p.gamma_A + p.R_AI), p.R_IA], [p.R_AI, -(p.gamma_I + p.R_IA)]])
    nA, nI = np.linalg.solve(A, [0.0, -P])
    return (nA, nI)


@pytest.mark.parametrize("kind,thr", [("single", O.threshold_single), ("double", O.threshold_double)])
def test_threshold_makes_the_gain_vanish(kind, thr):
    p = O.Physics()
    state = _fixed_point_at_zero_density(kind, p, thr(p))
    d = O.reservoir_derivatives(kind, p, 0.0, thr(p), tuple(np.float64(s) for s in state))
    assert max(abs(x) for x in d) < 1e-15
    assert abs(state[0] - p.gamma_C / p.R) < 1e-12 * p.gamma_C / p.R
    assert abs(O.threshold_single(p) - 0.02075) < 1e-12


@pytest.mark.parametrize("kind", ["single", "double", "quadratic-double"])
def test_condensed_uniform_steady_state_is_a_fixed_point(kind):
    """Closed-form uniform steady state above threshold: the reservoir stops and |psi| is constant."""
    p = O.Physics(g_I=0.001)
    g = O.Grid(4, 4, 4.0, 4.0)
    nA = p.gamma_C / p.R
    if kind == "single":
        P = 3 * O.threshold_single(p)
        rho = P / p.gamma_C - p.gamma_R / p.R
        state = (nA,)
    elif kind == "double":
        P = 3 * O.threshold_double(p)
        nI = (P + p.R_AI * nA) / (p.gamma_I + p.R_IA)
        rho = (p.R_IA * nI / nA - p.gamma_A - p.R_AI) / p.R
        state = (nA, nI)
    else:
        nI = 1.5 * math.sqrt(p.gamma_R * nA / p.kappa)
        P = p.kappa * nI ** 2 + p.gamma_I * nI
        rho = (p.kappa * nI ** 2 - p.gamma_R * nA) / (p.R * nA)
        state = (nA, nI)
    assert rho > 0
    psi = np.full((1, 1) + g.shape, math.sqrt(rho) * np.exp(0.3j))
    res = tuple(np.full((1, 1) + g.shape, s) for s in state)
    pump = np.full((1, 1) + g.shape, P)
    dpsi, dres = O.fdm_rhs(kind, p, g, psi, res, pump)
    for d in dres:
        np.testing.assert_allclose(d, 0, atol=1e-15)
    nI_eff = state[1] if kind != "single" else 0.0
    mu = p.g_C * rho + p.g_R * nA + p.g_I * nI_eff
    np.testing.assert_allclose(dpsi, (-1j * mu / p.hbar) * psi, atol=1e-15)


def test_reservoir_rk2_zero_pump_decay_factor():
    """psi = 0, P = 0: one midpoint step multiplies nR by 1 - gR dt + (gR dt)^2 / 2 exactly."""
    p = O.Physics()
    dt = 0.7
    n0 = np.full((1, 1, 3, 3), 2.5)
    psi = np.zeros((1, 1, 3, 3), dtype=complex)
    (n1,) = O.reservoir_rk2("single", p, dt, psi, 0.0, (n0,))
    x = p.gamma_R * dt
    np.testing.assert_allclose(n1, n0 * (1 - x + x * x / 2), rtol=1e-15)
    # second order against exp(-gR t)
    errs = []
    for h in (0.5, 0.25):
        n = n0.copy()
        for _ in range(int(round(20 / h))):
            (n,) = O.reservoir_rk2("single", p, h, psi, 0.0, (n,))
        errs.append(abs(n[0, 0, 0, 0] - 2.5 * math.exp(-p.gamma_R * 20)))
    assert 3.5 < errs[0] / errs[1] < 4.5


def test_reservoir_clamp_keeps_densities_non_negative():
    p = O.Physics()
    psi = np.zeros((1, 1, 2, 2), dtype=complex)
    n0 = np.array([-1.0, 0.5, -1e-3, 2.0]).reshape(1, 1, 2, 2)
    (n1,) = O.reservoir_rk2("single", p, 0.5, psi, 0.0, (n0,))
    x = p.gamma_R * 0.5
    np.testing.assert_array_equal(n1 == 0, n0 < 0)
    np.testing.assert_allclose(n1[n0 > 0], n0[n0 > 0] * (1 - x + x * x / 2), rtol=1e-15)


# ---------------------------------------------------------------- adaptive Lawson + controller
@pytest.mark.parametrize("tableau", ["dp5", "dop853"])
def test_lawson_adaptive_uniform_decay_closed_form(tableau):
    """Error <= 50 rtol against the closed form; a tighter rtol takes more accepted steps."""
    g = O.Grid(4, 4, 4.0, 4.0)
    p = O.Physics(g_C=0.5, gamma_C=0.4)
    accepted = []
    for rtol in (1e-6, 1e-9):
        prob = O.LawsonProblem("none", p, g, tableau=tableau, adaptive=True, rtol=rtol, atol=rtol * 1e-3)
        psi = np.ones((1, 1) + g.shape, dtype=complex)
        psi, _, ctl = O.lawson_advance_adaptive(prob, psi, (), _zero_pump(g), 0.0, 1.0, 3)
        assert np.max(np.abs(psi - _uniform_exact(p, 3.0))) <= 50 * rtol
        accepted.append(ctl.accepted)
    assert accepted[0] < accepted[1]


def test_lawson_fixed_rk4_is_fourth_order_on_uniform_decay():
    g = O.Grid(4, 4, 4.0, 4.0)
    p = O.Physics(g_C=0.5, gamma_C=0.4)
    errs = []
    for h in (0.2, 0.1):
        prob = O.LawsonProblem("none", p, g, tableau="rk4")
        psi = np.ones((1, 1) + g.shape, dtype=complex)
        for s in range(int(round(2.0 / h))):
            _, _, psi, _ = prob.step(psi, (), _zero_pump(g), s * h, h)
        errs.append(np.max(np.abs(psi - _uniform_exact(p, 2.0))))
    assert 13 < errs[0] / errs[1] < 19


# ---------------------------------------------------------------- spinor split-step
def _structured(g, C, seed=5):
    rng = np.random.default_rng(seed)
    return 0.3 * (rng.standard_normal((1, C) + g.shape) + 1j * rng.standard_normal((1, C) + g.shape))


@pytest.mark.parametrize("kouter", [False, True])
@pytest.mark.parametrize("kind", ["single", "double", "quadratic-double"])
def test_uncoupled_spinor_split_step_is_two_scalar_steps(kind, kouter):
    """No TE-TM, Zeeman or cross-spin term and one reservoir per component: each component
    advances as the scalar split step with its own reservoir and pump P w_c."""
    g = O.Grid(32, 24, 24.0, 18.0)
    p = O.Physics()
    psi = _structured(g, 2)
    nf = 1 if kind == "single" else 2
    res = tuple(np.full((1, 2) + g.shape, 0.7) for _ in range(nf))
    env = 0.05 * O.gaussian_profile(g, 5.0)
    pump_at = lambda t: env * (1.0 + t)
    w = np.asarray([0.7, 0.3]).reshape(1, 2, 1, 1)
    a_psi, a_res = psi, res
    scal = [(psi[:, c:c + 1], tuple(r[:, c:c + 1] for r in res)) for c in range(2)]
    fn = O.split_step_kouter if kouter else O.split_step_cpu
    for s in range(4):
        a_psi, a_res = O.split_step_spinor(kind, p, g, a_psi, a_res, pump_at, s * 0.01, 0.01, s, w, 2, kouter=kouter)
        scal = [fn(kind, p, g, ps, rs, (lambda t, c=c: pump_at(t) * w[:, c:c + 1]), s * 0.01, 0.01, s)
                for c, (ps, rs) in enumerate(scal)]
    for c in range(2):
        np.testing.assert_allclose(a_psi[:, c:c + 1], scal[c][0], rtol=1e-12, atol=1e-14)
        for x, y in zip(a_res, scal[c][1]):
            np.testing.assert_allclose(x[:, c:c + 1], y, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("kouter", [False, True])
def test_uniform_spinor_split_step_matches_closed_form(kouter):
    """Uniform psi+ / psi- without reservoir: |psi_c|^2 = |psi_c0|^2 e^{-gamma t} and the phase
    -(g rho_c0 + alpha_2 rho_o0)(1 - e^{-gamma t}) / (gamma hbar) -+ Z t / (2 hbar) solve the local
    ODE exactly, and the k = 0 propagator adds only the Zeeman phase: the step is exact."""
    g = O.Grid(8, 8, 8.0, 8.0)
    p = O.Physics(g_C=2e-3, alpha_2=-8e-4, zeeman=0.03, te_tm=0.05, gamma_C=0.3)
    a0, b0 = 0.8 + 0.3j, 0.4 - 0.5j
    psi = np.empty((1, 2) + g.shape, dtype=complex)
    psi[:, 0], psi[:, 1] = a0, b0
    dt, n = 0.05, 12
    for s in range(n):
        psi, _ = O.split_step_spinor("none", p, g, psi, (), lambda t: 0.0, s * dt, dt, s, None, 2, kouter=kouter)
    t = n * dt
    decay = (1 - np.exp(-p.gamma_C * t)) / p.gamma_C
    for c, (z0, o0, sgn) in enumerate(((a0, b0, -1), (b0, a0, 1))):
        phase = -(p.g_C * abs(z0) ** 2 + p.alpha_2 * abs(o0) ** 2) * decay / p.hbar + sgn * p.zeeman * t / (2 * p.hbar)
        exact = z0 * np.exp(-p.gamma_C * t / 2) * np.exp(1j * phase)
        np.testing.assert_allclose(psi[0, c], exact, rtol=1e-12)


def test_linear_spinor_split_step_is_the_te_tm_exponential():
    """Without nonlinearity, gain or reservoir the step is exp(-i H(k) dt / hbar) per k: a
    plane wave of psi- over five steps against scipy's expm of the 2 x 2 Hamiltonian."""
    g = O.Grid(16, 16, 16.0, 16.0)
    p = O.Physics(g_C=0.0, alpha_2=0.0, gamma_C=0.0, zeeman=0.02, te_tm=0.05)
    kx, ky = 2 * np.pi * 3 / g.lx, 2 * np.pi * 1 / g.ly
    wave = np.exp(1j * (kx * g.X + ky * g.Y))
    psi = np.zeros((1, 2) + g.shape, dtype=complex)
    psi[0, 1] = wave
    dt, n = 0.1, 5
    for s in range(n):
        psi, _ = O.split_step_spinor("none", p, g, psi, (), lambda t: 0.0, s * dt, dt, s, None, 2)
    E = p.hbar ** 2 * (kx * kx + ky * ky) / (2 * p.m_eff)
    H = np.array([[E + p.zeeman / 2, p.te_tm * (kx - 1j * ky) ** 2],
                  [p.te_tm * (kx + 1j * ky) ** 2, E - p.zeeman / 2]])
    amp = expm(-1j * H * n * dt / p.hbar) @ np.array([0.0, 1.0])
    np.testing.assert_allclose(psi[0, 0], amp[0] * wave, atol=1e-12)
    np.testing.assert_allclose(psi[0, 1], amp[1] * wave, atol=1e-12)


@pytest.mark.parametrize("kind", ["single", "double"])
def test_shared_spinor_reservoir_sees_the_total_density(kind):
    """One shared reservoir and psi- = 0 (no coupling): psi+ and the reservoir advance as the
    scalar split step; with psi- = psi+ the reservoir half-step sees twice the density."""
    g = O.Grid(16, 16, 16.0, 16.0)
    p = O.Physics()
    one = _structured(g, 1)
    nf = 1 if kind == "single" else 2
    res = tuple(np.full((1, 1) + g.shape, 0.5) for _ in range(nf))
    env = 0.05 * O.gaussian_profile(g, 4.0)
    pump_at = lambda t: env
    psi = np.concatenate((one, np.zeros_like(one)), axis=1)
    a_psi, a_res = O.split_step_spinor(kind, p, g, psi, res, pump_at, 0.0, 0.01, 0, None, 1)
    b_psi, b_res = O.split_step_cpu(kind, p, g, one, res, pump_at, 0.0, 0.01, 0)
    np.testing.assert_allclose(a_psi[:, :1], b_psi, rtol=1e-12, atol=1e-14)
    assert np.all(a_psi[:, 1] == 0)
    for x, y in zip(a_res, b_res):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-14)
    both = np.concatenate((one, one), axis=1)
    c_res = O._spinor_reservoir_half(kind, p, 0.005, both, env, res, None, 1)
    d_res = O.reservoir_rk2(kind, p, 0.005, np.sqrt(2.0) * one, env, res)
    for x, y in zip(c_res, d_res):
        np.testing.assert_allclose(x, y, rtol=1e-13)


@pytest.mark.parametrize("kind", ["none", "single", "double", "quadratic-double"])
@pytest.mark.parametrize("spinor", [False, True])
def test_ip_rk4_equals_lawson_rk4(kind, spinor):
    """The interaction-picture RK4 is Lawson's RK4 written with exp(+-L h/2) factors: on a
    structured field with a Gaussian pump, kinetic relaxation and (spinor) TE-TM / Zeeman /
    cross-spin terms, five steps agree to rounding, while a wrong stage propagator (exp(L h)
    at the midpoint) is caught by the same comparison."""
    g = O.Grid(32, 24, 24.0, 18.0)
    p = O.Physics(kinetic_relaxation_eta=0.03, te_tm=0.05 if spinor else 0.0, zeeman=0.02 if spinor else 0.0,
                  alpha_2=-1e-4 if spinor else 0.0)
    C = 2 if spinor else 1
    rng = np.random.default_rng(4)
    psi = (0.3 * (rng.standard_normal((1, C) + g.shape) + 1j * rng.standard_normal((1, C) + g.shape)))
    nf = {"none": 0, "single": 1, "double": 2, "quadratic-double": 2}[kind]
    res = tuple(rng.uniform(0.0, 2.0, (1, C) + g.shape) for _ in range(nf))
    env = 0.05 * O.gaussian_profile(g, 5.0)
    pump_at = lambda t: env * (1.0 + 0.5 * t)
    w = np.asarray([0.65, 0.35]).reshape(1, 2, 1, 1) if spinor else None
    prob = O.LawsonProblem(kind, p, g, tableau="rk4", pump_weights=w, reservoir_components=C)
    a_psi, a_res = psi, res
    b_psi, b_res = psi, res
    dt = 0.02
    for s in range(5):
        _, _, a_psi, a_res = prob.step(a_psi, a_res, pump_at, s * dt, dt)
        b_psi, b_res = O.ip_rk4_step(prob, b_psi, b_res, pump_at, s * dt, dt)
    assert O.rel_l2(b_psi, a_psi) <= 1e-13
    for x, y in zip(b_res, a_res):
        assert O.rel_l2(x, y) <= 1e-13
    # a planted mistake: the midpoint stages propagated by a full step
    bad = O.LawsonProblem(kind, p, g, tableau="rk4", pump_weights=w, reservoir_components=C)
    lin = bad.linear
    bad.linear = lambda fraction, h, fk: lin(1.0 if fraction == 0.5 else fraction, h, fk)
    _, _, c_psi, _ = bad.step(psi, res, pump_at, 0.0, dt)
    d_psi, _ = O.ip_rk4_step(prob, psi, res, pump_at, 0.0, dt)
    assert O.rel_l2(d_psi, c_psi) > 1e-6


def test_controller_rejects_then_lands_on_the_tick():
    g = O.Grid(4, 4, 4.0, 4.0)
    p = O.Physics(g_C=0.5, gamma_C=0.4)
    prob = O.LawsonProblem("none", p, g, tableau="dp5", adaptive=True, rtol=1e-6, atol=1e-9)
    psi = np.ones((1, 1) + g.shape, dtype=complex)
    psi, _, ctl = O.lawson_advance_adaptive(prob, psi, (), _zero_pump(g), 0.0, 5.0, 1)
    assert ctl.rejected > 0 and ctl.accepted > 1
    assert np.max(np.abs(psi - _uniform_exact(p, 5.0))) <= 50e-6


def test_controller_underflow_guard():
    ctl = O.StepController(1.0, 5, 4)
    ctl.h_next = 1e-300
    with pytest.raises(RuntimeError, match="underflow"):
        ctl.propose(1.0, 2.0)


def test_controller_factor_bounds():
    ctl = O.StepController(1.0, 5, 4)
    h, clamped, prop = ctl.propose(0.0, 10.0)
    ctl.update(True, 0.0, h, clamped, prop)
    assert ctl.h_next == 5.0                     # error 0 -> factor 5
    h, clamped, prop = ctl.propose(0.0, 100.0)
    ctl.update(False, 1e12, h, clamped, prop)
    assert ctl.h_next == pytest.approx(0.2 * h)  # huge error -> factor 0.2
    h2, clamped, prop = ctl.propose(0.0, 100.0)
    ctl.update(True, 1e-12, h2, clamped, prop)
    assert ctl.h_next <= h2                      # no growth right after a reject


# ---------------------------------------------------------------- steady state, initial field, pulses
def test_steady_metrics_of_scaled_field():
    rng = np.random.default_rng(1)
    psi = rng.standard_normal((2, 1, 8, 8)) + 1j * rng.standard_normal((2, 1, 8, 8))
    dens, norms, _, _ = O.steady_metrics(psi, None, 0.5)
    np.testing.assert_allclose(norms, 0.5 * (np.abs(psi) ** 2).sum(axis=(1, 2, 3)))
    _, _, rd, rn = O.steady_metrics(psi, dens, 0.5)
    np.testing.assert_array_equal(rd, 0)
    np.testing.assert_array_equal(rn, 0)
    s = np.array([1.1, 0.9]).reshape(2, 1, 1, 1)
    _, _, rd, rn = O.steady_metrics(psi * s, dens, 0.5)
    np.testing.assert_allclose(rd, np.abs(s.ravel() ** 2 - 1), rtol=1e-12)
    np.testing.assert_allclose(rn, np.abs(s.ravel() ** 2 - 1), rtol=1e-12)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_initial_psi_zero_mean_and_rms(precision):
    psi = O.initial_psi((32, 24), 1e-3, 7, precision)
    tol = 1e-6 if precision == "single" else 1e-13
    assert abs(np.mean(psi)) < tol * 1e-3
    assert abs(math.sqrt(np.mean(np.abs(psi) ** 2)) - 1e-3) < tol * 1e-3
    np.testing.assert_array_equal(psi, O.initial_psi((32, 24), 1e-3, 7, precision))
    assert not np.array_equal(psi, O.initial_psi((32, 24), 1e-3, 8, precision))


def test_pulse_train_peaks_and_cutoffs():
    kw = dict(P0=0.4, Pmax=1.2, sigma_time=0.1, cutoff_sigma=3.0, pulse_separation=0.6, n_pulses=0, delay=0.05)
    first = kw["delay"] + 0.3
    assert O.pulse_gaussian_coeff(first, **kw) == pytest.approx(0.4)
    assert O.pulse_gaussian_coeff(first + 0.6, **kw) == pytest.approx(1.2)
    assert O.pulse_gaussian_coeff(first + 1.2, **kw) == pytest.approx(1.2)
    assert O.pulse_gaussian_coeff(first + 0.1, **kw) == pytest.approx(0.4 * math.exp(-0.5))
    assert O.pulse_gaussian_coeff(0.0, **kw) == 0.0
    sparse = {**kw, "pulse_separation": 1.0}
    assert O.pulse_gaussian_coeff(first + 0.29, **sparse) > 0.0
    assert O.pulse_gaussian_coeff(first + 0.31, **sparse) == 0.0          # beyond 3 sigma, before the next pulse
    assert O.pulse_gaussian_coeff(first + 0.6, **{**kw, "n_pulses": 1}) == 0.0


# ---------------------------------------------------------------- closed interval and 9-point stencils
def _sym5(k, h):
    return -(2 - 2 * np.cos(k * h)) / h ** 2


def _sym9(kx, ky, h):
    a, b = kx * h, ky * h
    return (4 * (2 * np.cos(a) + 2 * np.cos(b)) + 4 * np.cos(a) * np.cos(b) - 20) / (6 * h ** 2)


@pytest.mark.parametrize("kind", ["five-point", "isotropic-9pt"])
def test_closed_interval_neumann_cosines_are_eigenfunctions(kind):
    """cos(m pi (x + L/2) / L) on the nodes is even about both ends, so the mirrored stencil keeps it."""
    g = O.Grid(33, 33, 16.0, 16.0, boundary="closed-interval")
    assert g.dx == 0.5 and g.X[0, 0] == -8.0 and g.X[0, -1] == 8.0
    kx, ky = 3 * np.pi / g.lx, 2 * np.pi / g.ly
    f = np.cos(kx * (g.X + g.lx / 2)) * np.cos(ky * (g.Y + g.ly / 2))
    lap = O.laplacian(f, g.dx, g.dy, "closed-interval", kind)
    sym = _sym5(kx, g.dx) + _sym5(ky, g.dy) if kind == "five-point" else _sym9(kx, ky, g.dx)
    np.testing.assert_allclose(lap, sym * f, atol=1e-12 * abs(sym))
    # the constant is in the kernel (Neumann): Laplacian exactly zero
    np.testing.assert_array_equal(O.laplacian(np.full(g.shape, 2.5), g.dx, g.dy, "closed-interval", kind), 0.0)


def test_nine_point_symbol_and_isotropy():
    """Periodic 9-point stencil on plane cosines gives its symbol; its O(h^2) error is isotropic."""
    g = O.Grid(48, 48, 24.0, 24.0)
    h = g.dx
    kx, ky = 2 * np.pi * 5 / g.lx, 2 * np.pi * 3 / g.ly
    f = np.cos(kx * g.X) * np.cos(ky * g.Y)
    np.testing.assert_allclose(O.laplacian(f, h, h, "periodic", "isotropic-9pt"), _sym9(kx, ky, h) * f,
                               atol=1e-12 * abs(_sym9(kx, ky, h)))
    k = 0.6
    along = _sym9(k, 0.0, h) + k ** 2
    diag = _sym9(k / np.sqrt(2), k / np.sqrt(2), h) + k ** 2
    five_along = _sym5(k, h) + k ** 2
    five_diag = 2 * _sym5(k / np.sqrt(2), h) + k ** 2
    assert abs(along - diag) < 0.01 * abs(along)                 # 9-point: direction-free at O(h^2)
    assert abs(five_along - five_diag) > 0.3 * abs(five_along)    # 5-point: not


def test_closed_grid_matches_reference_geometry():
    g = O.Grid(5, 3, 4.0, 2.0, boundary="closed-interval")
    assert (g.dx, g.dy) == (1.0, 1.0)
    np.testing.assert_array_equal(g.X[0], [-2, -1, 0, 1, 2])
    np.testing.assert_array_equal(g.Y[:, 0], [-1, 0, 1])


# ---------------------------------------------------------------- kinetic, eta and TE-TM terms of the FDM step
def _rk4_poly(z):
    """Stability polynomial of classic RK4: one step of y' = lambda y multiplies y by R(h lambda)."""
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def _rk4_poly_matrix(Z):
    eye = np.eye(Z.shape[0], dtype=complex)
    return eye + Z + Z @ Z / 2 + Z @ Z @ Z / 6 + Z @ Z @ Z @ Z / 24


def _linear_physics():
    """No self-interaction and no stimulated scattering: the psi equation is linear and the
    reservoir, held at its fixed point nA0 by P = gamma_R nA0, acts as a constant."""
    return O.Physics(g_C=0.0, R=0.0, gamma_C=0.07, g_R=0.002, gamma_R=0.005, kinetic_relaxation_eta=0.3)


def _linear_rate(p, lam, nA0):
    """d psi / dt = rate psi for a Laplacian eigenfunction with eigenvalue -lam (lam >= 0):
    kinetic -(hbar^2/2m) lap gives -i (hbar/2m) lam, g_R nA0 a phase, -gamma_C/2 the loss and
    eta nA0 (hbar/2m) lap the damping -eta nA0 (hbar/2m) lam."""
    hb2m = p.hbar / (2 * p.m_eff)
    return -1j * (hb2m * lam + p.g_R * nA0 / p.hbar) - p.gamma_C / 2 - p.kinetic_relaxation_eta * nA0 * hb2m * lam


@pytest.mark.parametrize("stencil", ["five-point", "isotropic-9pt"])
def test_rk4_fdm_plane_wave_follows_the_discrete_symbol(stencil):
    """A plane wave stays a plane wave: after n steps psi = R(dt rate)^n psi0, with rate from the
    stencil's symbol.  A flipped kinetic sign, a wrong hbar/2m or a wrong eta term fails this."""
    g = O.Grid(48, 48, 36.0, 36.0)
    p = _linear_physics()
    nA0, dt, n = 1.7, 0.05, 12
    kx, ky = 2 * np.pi * 3 / g.lx, -2 * np.pi * 2 / g.ly
    psi = (0.2 * np.exp(1j * (kx * g.X + ky * g.Y))).reshape(1, 1, *g.shape)
    res = (np.full((1, 1) + g.shape, nA0),)
    pump = np.full((1, 1) + g.shape, p.gamma_R * nA0)
    sym = _sym5(kx, g.dx) + _sym5(ky, g.dy) if stencil == "five-point" else _sym9(kx, ky, g.dx)
    z = dt * _linear_rate(p, -sym, nA0)
    expect = _rk4_poly(z) ** n * psi
    for s in range(n):
        psi, res = O.rk4_fdm_step("single", p, g, psi, res, lambda t: pump, s * dt, dt, laplacian_kind=stencil)
    np.testing.assert_allclose(psi, expect, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(res[0], nA0)
    # the phase turns the way exp(-i E t / hbar) does: clockwise for positive kinetic energy
    assert np.angle(_rk4_poly(z)) < 0


@pytest.mark.parametrize("stencil", ["five-point", "isotropic-9pt"])
def test_rk4_fdm_neumann_cosine_on_a_closed_interval(stencil):
    """On the closed interval the Neumann cosine is an eigenfunction of the mirrored stencil:
    psi_n = R(dt rate)^n psi_0 as for the plane wave."""
    g = O.Grid(33, 25, 16.0, 12.0, boundary="closed-interval")
    assert g.dx == g.dy == 0.5
    p = _linear_physics()
    nA0, dt, n = 0.9, 0.05, 10
    kx, ky = 3 * np.pi / g.lx, 2 * np.pi / g.ly
    psi = (0.3 * np.cos(kx * (g.X + g.lx / 2)) * np.cos(ky * (g.Y + g.ly / 2))).astype(complex).reshape(1, 1, *g.shape)
    res = (np.full((1, 1) + g.shape, nA0),)
    pump = np.full((1, 1) + g.shape, p.gamma_R * nA0)
    sym = _sym5(kx, g.dx) + _sym5(ky, g.dy) if stencil == "five-point" else _sym9(kx, ky, g.dx)
    expect = _rk4_poly(dt * _linear_rate(p, -sym, nA0)) ** n * psi
    for s in range(n):
        psi, res = O.rk4_fdm_step("single", p, g, psi, res, lambda t: pump, s * dt, dt, laplacian_kind=stencil)
    np.testing.assert_allclose(psi, expect, rtol=1e-12, atol=1e-14)


def test_rk4_fdm_spinor_plane_wave_follows_the_discrete_te_tm_matrix():
    """Both components on one plane wave: d psi/dt = (-i/hbar) H_d psi - gamma_C/2 psi with the
    stencil's 2x2 matrix H_d = [[E + Z/2, T_+], [T_-, E - Z/2]], T_-/+ = -beta[(sx - sy) -/+ 2i sxy]
    from the second differences and the 4-corner mixed difference.  n RK4 steps multiply the
    spinor by R(dt M)^n.  This fixes the TE-TM sign and which component each row couples to, and
    the Zeeman sign, in the whole step."""
    g = O.Grid(40, 32, 40.0, 32.0)
    p = O.Physics(g_C=0.0, R=0.0, gamma_C=0.05, te_tm=0.05, zeeman=0.02, alpha_2=0.0)
    dt, n = 0.1, 8
    kx, ky = 2 * np.pi * 2 / g.lx, 2 * np.pi * 3 / g.ly
    wave = np.exp(1j * (kx * g.X + ky * g.Y))
    a = np.array([0.3 + 0.1j, -0.2 + 0.25j])
    psi = (a[:, None, None] * wave)[None]
    sx, sy = _sym5(kx, g.dx), _sym5(ky, g.dy)
    sxy = -np.sin(kx * g.dx) * np.sin(ky * g.dy) / (g.dx * g.dy)
    E = -(p.hbar ** 2) / (2 * p.m_eff) * (sx + sy)
    H = np.array([[E + p.zeeman / 2, -p.te_tm * ((sx - sy) - 2j * sxy)],
                  [-p.te_tm * ((sx - sy) + 2j * sxy), E - p.zeeman / 2]])
    M = (-1j / p.hbar) * H - p.gamma_C / 2 * np.eye(2)
    amp = np.linalg.matrix_power(_rk4_poly_matrix(dt * M), n) @ a
    expect = (amp[:, None, None] * wave)[None]
    for s in range(n):
        psi, _ = O.rk4_fdm_step("none", p, g, psi, (), _zero_pump(g), s * dt, dt)
    np.testing.assert_allclose(psi, expect, rtol=1e-12, atol=1e-14)
    # the continuum limit of H_d is the dispersion matrix: T_+ -> beta (kx - i ky)^2
    Hc = O.spinor_hamiltonian(g, p)[:, :, 3, 2]
    assert abs(H[0, 1] - Hc[0, 1]) < 0.1 * abs(Hc[0, 1])       # O((k dx)^2) apart; a sign error is O(1)
    assert abs(H[1, 0] - Hc[1, 0]) < 0.1 * abs(Hc[1, 0])


def test_rk4_fdm_uniform_spinor_alpha2_and_zeeman_closed_form():
    """Without gain or loss each uniform component only turns its phase:
    psi_+/-(t) = psi_+/-(0) exp(-i (g_C |psi_+/-|^2 + alpha_2 |psi_-/+|^2 +/- Z/2) t / hbar).
    RK4 converges to it at fourth order; alpha_2 on the wrong density or a flipped Zeeman sign
    leaves an O(1) error."""
    g = O.Grid(4, 4, 4.0, 4.0)
    p = O.Physics(g_C=0.4, gamma_C=0.0, R=0.0, alpha_2=-0.15, zeeman=0.05)
    a = np.array([0.9 + 0.2j, 0.35 - 0.1j])
    rho = np.abs(a) ** 2
    T = 2.0
    mu = np.array([p.g_C * rho[0] + p.alpha_2 * rho[1] + p.zeeman / 2,
                   p.g_C * rho[1] + p.alpha_2 * rho[0] - p.zeeman / 2])
    exact = a * np.exp(-1j * mu * T / p.hbar)
    errs = []
    for dt in (0.2, 0.1):
        psi = np.broadcast_to(a[:, None, None], (2,) + g.shape)[None].astype(complex)
        for s in range(int(round(T / dt))):
            psi, _ = O.rk4_fdm_step("none", p, g, psi, (), _zero_pump(g), s * dt, dt)
        errs.append(np.max(np.abs(psi[0, :, 0, 0] - exact)))
    assert errs[1] < 1e-6
    assert 13 < errs[0] / errs[1] < 19


# ---------------------------------------------------------------- embedded error norm
def _increment(tab, weights, z):
    """z w^T (I - z A)^-1 1: one step's increment of y' = lambda y per unit y, z = h lambda."""
    s = len(tab.c)
    stage = np.linalg.solve(np.eye(s) - z * tab.A, np.ones(s, dtype=complex))
    return z * np.dot(weights, stage)


def _stability(tab, weights, z):
    """1 + z w^T (I - z A)^-1 1 for the tableau's A and the weight vector w."""
    return 1 + _increment(tab, weights, z)


@pytest.mark.parametrize("tableau", ["dp5", "dop853"])
def test_lawson_error_norm_closed_form_two_members_with_reservoir(tableau):
    """Linear uniform problem: psi' = mu psi (mu = -i V/hbar - gamma_C/2) and, decoupled,
    n' = P - gamma_R n.  The embedded error of each field is (R_b - R_bhat)(z) times its
    deviation, so the per-member norm follows from the stability functions alone:
    DP5 sqrt(S5 / count), DOP853 S5 / sqrt((S5 + 0.01 S3) count), S = sum of squared scaled
    errors over psi and reservoir points, count = both point counts; the trial reports the
    largest member.  A dropped count, a wrong blend or a missing member max fails."""
    g = O.Grid(4, 4, 4.0, 4.0)
    V = 0.05
    p = O.Physics(g_C=0.0, R=0.0, g_R=0.0, gamma_C=0.6, gamma_R=0.3)
    tab = O.oracle_tableau(tableau)
    h, rtol, atol, P0 = (0.8 if tableau == "dp5" else 2.0), 1e-6, 1e-3, 0.02
    a = np.array([0.04 - 0.02j, 1.0 + 0.5j])
    n0 = np.array([0.01, 0.2])
    psi = np.broadcast_to(a[:, None, None, None], (2, 1) + g.shape).astype(complex)
    res = (np.broadcast_to(n0[:, None, None, None], (2, 1) + g.shape).astype(float),)
    pump = np.full((1, 1) + g.shape, P0)
    prob = O.LawsonProblem("single", p, g, potential=V, tableau=tableau, adaptive=True, rtol=rtol, atol=atol)
    accepted, err, psi1, res1 = prob.step(psi, res, lambda t: pump, 0.0, h)

    z = h * (-1j * V / p.hbar - p.gamma_C / 2)
    zr = -p.gamma_R * h
    nstar = P0 / p.gamma_R
    b = np.asarray(tab.b, dtype=float)
    np.testing.assert_allclose(psi1[:, 0, 0, 0], _stability(tab, b, z) * a, rtol=1e-13)
    np.testing.assert_allclose(res1[0][:, 0, 0, 0], nstar + _stability(tab, b, zr).real * (n0 - nstar), rtol=1e-13)

    def sums(bhat):
        w = b - np.asarray(bhat, dtype=float)
        e_psi = np.abs(_increment(tab, w, z) * a)
        e_res = np.abs(_increment(tab, w, zr).real * (n0 - nstar))
        sc_psi = atol + rtol * np.maximum(np.abs(a), np.abs(psi1[:, 0, 0, 0]))
        sc_res = atol + rtol * np.maximum(np.abs(n0), np.abs(res1[0][:, 0, 0, 0]))
        return 16 * (e_psi / sc_psi) ** 2 + 16 * (e_res / sc_res) ** 2

    count = 32
    s5 = sums(tab.b_hat)
    if tab.b_hat3 is None:
        member = np.sqrt(s5 / count)
    else:
        s3 = sums(tab.b_hat3)
        member = s5 / np.sqrt((s5 + 0.01 * s3) * count)
    assert member[1] > 1.5 * member[0]            # the second member decides
    if tab.b_hat3 is not None:
        assert np.all(0.01 * s3 > 0.1 * s5)       # the blend term matters
    # the oracle sums the embedded terms field by field, so it carries their cancellation
    assert math.isclose(err, float(member.max()), rel_tol=1e-7)
    assert accepted == (err <= 1)


# ---------------------------------------------------------------- Wigner noise in the Lawson step
@pytest.mark.parametrize("precision", ["single", "double"])
def test_lawson_noise_is_the_wigner_increment_on_the_final_state(precision):
    """The noisy fixed step equals the noiseless one plus the Wigner increment of that step's
    index, scaled by the final reservoir; the reservoir itself is not touched by the noise."""
    g = O.Grid(16, 16, 20.0, 20.0)
    p = O.Physics()
    real = np.float32 if precision == "single" else np.float64
    rng = np.random.default_rng(11)
    psi = (0.2 * (rng.standard_normal((1, 1) + g.shape) + 1j * rng.standard_normal((1, 1) + g.shape))).astype(
        np.complex64 if precision == "single" else np.complex128)
    res = (np.abs(rng.standard_normal((1, 1) + g.shape)).astype(real),)
    pump = np.full((1, 1) + g.shape, 0.05, dtype=real)
    prob = O.LawsonProblem("single", p, g, precision=precision, tableau="rk4")
    noise = dict(seeds=[1234], gamma=[p.gamma_C], scattering=[p.R], strength=[1.0], volume=g.dx * g.dy,
                 precision=precision)
    _, _, plain, res_plain = prob.step(psi, res, lambda t: pump, 0.3, 0.01)
    _, _, noisy, res_noisy = prob.step(psi, res, lambda t: pump, 0.3, 0.01, noise=noise, step=7)
    np.testing.assert_array_equal(res_noisy[0], res_plain[0])
    inc = np.zeros_like(plain)
    O.wigner_add(inc, res_plain[0], dt=0.01, step=7, **noise)
    np.testing.assert_array_equal(noisy - plain, (plain + inc) - plain)
    sigma = np.sqrt(p.gamma_C + p.R * res_plain[0].astype(np.float64)) * np.sqrt(0.01 / (4 * g.dx * g.dy))
    z = (inc / sigma)[0, 0]
    assert abs(np.mean(z.real)) < 0.3 and 0.6 < np.var(z.real) < 1.5 and 0.6 < np.var(z.imag) < 1.5


# ---------------------------------------------------------------- the package's PI controller
def test_package_controller_follows_the_oracle_controller():
    """polarism_amd.adaptive.PIControl (host half of the device controller) and the oracle's
    StepController give the same proposals for the same error sequence, rejections,
    clamped landings and a dt_max included."""
    import random
    from polarism_amd.adaptive import PIControl
    rnd = random.Random(3)
    for dt_max in (None, 0.7):
        pkg = PIControl(k=5.0, h_next=min(1.0, dt_max or 1.0), dt_max=dt_max)
        ora = O.StepController(1.0, 5, 4, dt_max=dt_max)
        t, target = 0.0, 3.3
        for _ in range(60):
            if t >= target:
                target += 1.7
            h1, c1, p1 = pkg.propose(t, target)
            h2, c2, p2 = ora.propose(t, target)
            assert (h1, c1, p1) == (h2, c2, p2)
            err = rnd.choice([0.0, 1e-6, 0.3, 0.99, 1.0, 1.01, 4.0, 80.0, float("nan")])
            ok = err <= 1
            pkg.settle(h1, c1, p1, ok, err)
            ora.update(ok, err, h2, c2, p2)
            assert pkg.h_next == ora.h_next and pkg.no_growth == ora.no_growth
            if ok:
                t = target if c1 else t + h1
        assert (pkg.accepted, pkg.rejected) == (ora.accepted, ora.rejected)
