"""CPU oracle for the driven-dissipative Gross-Pitaevskii time stepper.

TEST INFRASTRUCTURE ONLY.  Nothing in ``polarism_amd`` imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker or as the timed CPU baseline.

It restates, in plain NumPy, the reference's CPU algorithms for the hot path
(file:line citations are relative to the Polarism 0.9.2 checkout):

* Philox4x32-10 words, counters and Box-Muller normals -- noise/philox.py:7-56
* Wigner increment -- noise/stochastic.py:224-281
* initial condensate field -- noise/initial_field.py:363-380
* periodic grid -- grid/periodic_grid.py:35-69
* reservoir derivatives -- reservoir/single_reservoir.py:81-111,
  reservoir/double_reservoir.py:255-292,
  reservoir/quadratic_double_reservoir.py:131-226
* scalar and spinor RHS with the 5-point Laplacian -- model.py:89-248
* RK4-FDM step -- solver/rk4_fdm_solver.py:36-77
* split-step FFT, CPU order -- solver/split_step_fft_solver.py:54-159
* split-step FFT, K-outer order -- tests/crosscheck/test_split_step_fft_cuda.py:46-70
* spinor split-step -- solver/split_step_fft_solver.py:46-53, 81-94, 124-145;
  K-outer spinor order -- solver/split_step_fft_cuda_solver.py:229-330
* Lawson IFRK generic step + embedded error -- solver/ifrk_fft_cuda_solver.py:96-293
* interaction-picture RK4 -- solver/ip_rk4_solver.py:122-213
* PI step controller -- solver/step_controller.py:48-81
* spinor dispersion and 2x2 exponential -- dispersion.py:34-89
* steady-state test -- simulation_controller.py:710-734

Pinning (what checks each function; tests/ file in brackets):
* philox4x32: Random123 known-answer vectors held by the reference's
  tests/unit/test_noise.py:56-64 [test_oracle_noise.py].
* uniforms / normal_pair: bit-exact uniforms from the closed form; normals vs
  math.log/cos/sin; moments; one-step Wigner variance [test_oracle_noise.py].
* LawsonProblem.step (rk4 tableau; single and quadratic-double reservoirs;
  eta 0 and 0.04; complex potential; pulsed pump; f32/f64): the reference's
  golden trajectories tests/data/ifrk4_reference/reference.npz
  [test_oracle_golden.py]; 4th order on the uniform closed form [test_oracle_pins.py].
* ip_rk4_step: equals LawsonProblem.step with the rk4 tableau (the same method written
  with exp(+-L h/2) factors) to rounding, scalar and spinor, every reservoir kind
  [test_oracle_pins.py]; through it, the golden trajectories above.
* LawsonProblem adaptive + error + StepController + lawson_advance_adaptive:
  closed-form uniform decay within 50 rtol (tests/unit/test_adaptive_dt.py:34-38),
  rejection / tick landing, factor bounds, underflow guard; the embedded error norm
  (DP5 sqrt(S/count), DOP853 blend, psi and reservoir points counted, max over two
  members) against the embedded stability functions of the tableaus on a linear
  uniform problem [test_oracle_pins.py].
* LawsonProblem.step with noise: the noisy step minus the plain one is exactly the
  Wigner increment of that step index with the final reservoir; the reservoir is
  untouched [test_oracle_pins.py].
* oracle_tableau: order conditions, c = A 1, and an independent rational
  transcription [test_tableaus.py].
* laplacian: discrete cosine eigenfunction, second order [test_oracle_pins.py].
* te_tm_energy: plane-wave stencil symbols and the continuum limit equal to
  the dispersion matrix off-diagonal [test_oracle_pins.py].
* fdm_rhs: uniform closed-form steady states of all three reservoirs (a fixed
  point with mu/hbar phase rotation), Zeeman sign vs a scalar potential
  [test_oracle_pins.py].
* rk4_fdm_step: 4th-order convergence on the uniform closed form; the kinetic,
  eta, g_R and loss terms by plane waves (periodic, 5- and 9-point symbols) and
  Neumann cosines (closed interval) after n steps, psi_n = R(dt rate)^n psi_0 with
  R the RK4 stability polynomial; the spinor TE-TM / Zeeman 2x2 stencil matrix on a
  plane wave (matrix polynomial); alpha_2 and Zeeman on a uniform spinor closed
  form [test_oracle_pins.py].
* reservoir_derivatives / reservoir_rk2 / thresholds: the threshold makes the
  gain vanish at the psi = 0 fixed point; exact RK2 decay factor; clamp.
* split_step_*: exact for uniform fields, plane-wave phase, orderings
  differing at O(dt^2) [test_oracle_pins.py].
* spinor_propagator: scipy.linalg.expm per k, unitarity [test_oracle_pins.py].
* steady_metrics, initial_psi, pulse_gaussian_coeff: invariants and closed
  forms [test_oracle_pins.py].
* lattice_sites / lattice_profile / supergaussian_profile: hand-listed square sites,
  nearest-neighbour distances and coordination numbers of the five lattices, defects,
  a 90-degree rotation of the square lattice, the disorder draws, super-Gaussian order 1
  equal to the Gaussian [test_pumps_host.py].
Parity unpinned (no reference execution possible; restatement-pinned only):
the full multi-step trajectories of rk4_fdm_step and split_step_* at the
quickstart / split-step configurations, and the exact DP5 / DOP853 trial
sequences of the reference.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

# ---------------------------------------------------------------------------
# Philox4x32-10 and Box-Muller (noise/philox.py:7-56)
# ---------------------------------------------------------------------------
_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = 0x9E3779B9
_W1 = 0xBB67AE85
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32(counter, key: int) -> np.ndarray:
    """Ten Philox rounds on ``(..., 4)`` uint32 counters with a 64-bit key."""
    ctr = np.asarray(counter, dtype=np.uint32)
    if ctr.shape[-1] != 4:
        raise ValueError("counter needs 4 words")
    x = [ctr[..., i].astype(np.uint64) for i in range(4)]
    k0 = int(key) & 0xFFFFFFFF
    k1 = (int(key) >> 32) & 0xFFFFFFFF
    for r in range(10):
        hi_lo0 = x[0] * _M0
        hi_lo1 = x[2] * _M1
        y0 = ((hi_lo1 >> np.uint64(32)) ^ x[1] ^ np.uint64(k0)) & _MASK32
        y1 = hi_lo1 & _MASK32
        y2 = ((hi_lo0 >> np.uint64(32)) ^ x[3] ^ np.uint64(k1)) & _MASK32
        y3 = hi_lo0 & _MASK32
        x = [y0, y1, y2, y3]
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return np.stack([v.astype(np.uint32) for v in x], axis=-1)


def counter_words(cells, step: int, stream: int = 1) -> np.ndarray:
    """Counter (cell, step low word, step high word, stream) (philox.py:30-38)."""
    cells = np.asarray(cells, dtype=np.uint32)
    out = np.empty(cells.shape + (4,), dtype=np.uint32)
    out[..., 0] = cells
    out[..., 1] = np.uint32(int(step) & 0xFFFFFFFF)
    out[..., 2] = np.uint32((int(step) >> 32) & 0xFFFFFFFF)
    out[..., 3] = np.uint32(stream)
    return out


def uniforms(words, precision: str):
    """The two uniforms in (0, 1] that Box-Muller consumes (philox.py:44-53)."""
    w = np.asarray(words, dtype=np.uint32)
    if precision == "single":
        u1 = ((w[..., 0] >> 8).astype(np.float32) + np.float32(1)) * np.float32(2.0 ** -24)
        u2 = ((w[..., 1] >> 8).astype(np.float32) + np.float32(1)) * np.float32(2.0 ** -24)
        return u1, u2
    u1 = ((w[..., 0] >> 5).astype(np.float64) * 67108864.0
          + (w[..., 1] >> 6).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = ((w[..., 2] >> 5).astype(np.float64) * 67108864.0
          + (w[..., 3] >> 6).astype(np.float64) + 1.0) * 2.0 ** -53
    return u1, u2


def normal_pair(words, precision: str):
    """Box-Muller (re, im) standard normals from Philox words (philox.py:41-56)."""
    u1, u2 = uniforms(words, precision)
    if precision == "single":
        r = np.sqrt(np.float32(-2.0) * np.log(u1))
        a = np.float32(2.0 * np.pi) * u2
    else:
        r = np.sqrt(-2.0 * np.log(u1))
        a = (2.0 * np.pi) * u2
    return r * np.cos(a), r * np.sin(a)


def wigner_add(field, active, seeds, gamma, scattering, strength, dt, volume, step,
               precision):
    """Add one truncated-Wigner increment in place (noise/stochastic.py:224-281).

    ``field`` is ``(B, C, *S)`` complex, ``active`` broadcastable to
    ``(B, C_res, *S)`` (or 0).  Component c of a spinor uses stream 1 + c.
    """
    real = np.float32 if precision == "single" else np.float64
    spatial = field.shape[2:]
    n = int(np.prod(spatial))
    cells = np.arange(n)
    for b in range(field.shape[0]):
        if strength[b] == 0.0:
            continue
        for c in range(field.shape[1]):
            words = philox4x32(counter_words(cells, step, 1 + c), int(seeds[b]))
            re, im = normal_pair(words, precision)
            if np.ndim(active) == 0:
                dens = 0
            else:
                a = np.asarray(active)
                dens = a[b, c if a.shape[1] == 2 else 0].reshape(-1)
            sigma = np.sqrt(real(strength[b]) * (real(gamma[b]) + real(scattering[b]) * dens)
                            * real(dt / (4.0 * volume)))
            field[b, c] += (sigma * (re + 1j * im)).reshape(spatial)


# ---------------------------------------------------------------------------
# Initial field and grid
# ---------------------------------------------------------------------------
def initial_psi(shape, eps, seed, precision="double"):
    """Zero-mean complex Gaussian with RMS ``eps`` (initial_field.py:372-380)."""
    rdtype = np.float32 if precision == "single" else np.float64
    cdtype = np.complex64 if precision == "single" else np.complex128
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(shape).astype(rdtype)
    im = rng.standard_normal(shape).astype(rdtype)
    psi = re + 1j * im
    psi -= np.mean(psi)
    rms = float(np.sqrt(np.mean(np.abs(psi) ** 2)))
    if rms < 1e-30:
        rms = 1.0
    return ((eps / rms) * psi).astype(cdtype)


@dataclass
class Grid:
    """2-D grid, x the fastest axis.  Periodic (grid/periodic_grid.py:35-69): dx = lx/nx,
    centred X = (i - (n-1)/2) dx.  Closed interval (grid/closed_interval_grid.py:35-80):
    dx = lx/(nx-1), X = linspace(-lx/2, lx/2, nx)."""
    nx: int
    ny: int
    lx: float
    ly: float
    boundary: str = "periodic"

    def __post_init__(self):
        if self.boundary == "periodic":
            self.dx = self.lx / self.nx
            self.dy = self.ly / self.ny
            x = (np.arange(self.nx, dtype=np.float64) - (self.nx - 1) * 0.5) * self.dx
            y = (np.arange(self.ny, dtype=np.float64) - (self.ny - 1) * 0.5) * self.dy
        elif self.boundary == "closed-interval":
            self.dx = self.lx / (self.nx - 1)
            self.dy = self.ly / (self.ny - 1)
            x = np.linspace(-self.lx / 2, self.lx / 2, self.nx, dtype=np.float64)
            y = np.linspace(-self.ly / 2, self.ly / 2, self.ny, dtype=np.float64)
        else:
            raise ValueError(self.boundary)
        self.X, self.Y = np.meshgrid(x, y, indexing="xy")
        self.kx = 2 * np.pi * np.fft.fftfreq(self.nx, d=self.dx)
        self.ky = 2 * np.pi * np.fft.fftfreq(self.ny, d=self.dy)
        self.KX, self.KY = np.meshgrid(self.kx, self.ky, indexing="xy")
        self.k_squared = self.KX ** 2 + self.KY ** 2
        self.shape = (self.ny, self.nx)


def gaussian_profile(grid, sigma, x0=0.0, y0=0.0):
    """exp(-r^2/(2 sigma^2)) (laser/continuous_gaussian.py:39-41)."""
    r2 = (grid.X - x0) ** 2 + (grid.Y - y0) ** 2
    return np.exp(-0.5 * r2 / sigma ** 2)


def ring_profile(grid, radius, width, x0=0.0, y0=0.0):
    """Gaussian radial cross-section (profiles/builtin_profiles.py:66-74)."""
    r = np.sqrt((grid.X - x0) ** 2 + (grid.Y - y0) ** 2)
    return np.exp(-0.5 * ((r - radius) / width) ** 2)


def supergaussian_profile(grid, sigma, order, x0=0.0, y0=0.0):
    """exp(-0.5 (r^2 / sigma^2)^order) (profiles/builtin_profiles.py:45-56)."""
    r2 = (grid.X - x0) ** 2 + (grid.Y - y0) ** 2
    return np.exp(-0.5 * (r2 / sigma ** 2) ** order)


_LATTICE_GEOMETRY = {
    # primitive vectors and sublattice offsets in units of the constant (builtin_profiles.py:181-197)
    "square": ([[1, 0], [0, 1]], [[0, 0]]),
    "triangular": ([[1, 0], [0.5, np.sqrt(3) / 2]], [[0, 0]]),
    "honeycomb": ([[1, 0], [0.5, np.sqrt(3) / 2]], [[0, 0], [0.5, 0.5 / np.sqrt(3)]]),
    "kagome": ([[2, 0], [1, np.sqrt(3)]], [[0, 0], [1, 0], [0.5, np.sqrt(3) / 2]]),
    "lieb": ([[2, 0], [0, 2]], [[0, 0], [1, 0], [0, 1]]),
}


def lattice_sites(lx, ly, lattice, constant, rotation=0.0, x0=0.0, y0=0.0, defects=()):
    """(N, 5) rows (i, j, s, x, y) of the sites covering the box plus one cell, in i, j, s order
    (builtin_profiles.py:200-224): cell range from the box corners in cell coordinates +-2,
    kept where |x| <= lx/2 + m and |y| <= ly/2 + m, m = the longer primitive vector."""
    vec, offs = _LATTICE_GEOMETRY[lattice]
    c, s_ = np.cos(np.radians(rotation)), np.sin(np.radians(rotation))
    R = np.array([[c, -s_], [s_, c]])
    B = constant * np.asarray(vec, dtype=np.float64) @ R.T
    Off = constant * np.asarray(offs, dtype=np.float64) @ R.T
    o = np.array([x0, y0])
    corners = np.array([[-lx / 2, -ly / 2], [-lx / 2, ly / 2], [lx / 2, -ly / 2], [lx / 2, ly / 2]]) - o
    cc = corners @ np.linalg.inv(B)
    lo = np.floor(cc.min(axis=0)).astype(int) - 2
    hi = np.ceil(cc.max(axis=0)).astype(int) + 2
    I, J, S = np.meshgrid(np.arange(lo[0], hi[0] + 1), np.arange(lo[1], hi[1] + 1), np.arange(len(Off)),
                          indexing="ij")
    I, J, S = I.ravel(), J.ravel(), S.ravel()
    P = o + I[:, None] * B[0] + J[:, None] * B[1] + Off[S]
    m = max(np.hypot(*B[0]), np.hypot(*B[1]))
    keep = (np.abs(P[:, 0]) <= lx / 2 + m) & (np.abs(P[:, 1]) <= ly / 2 + m)
    for d in defects:
        keep &= ~((I == d[0]) & (J == d[1]) & (S == d[2]))
    return np.column_stack([I, J, S, P[:, 0], P[:, 1]])[keep]


def lattice_profile(grid, lattice, constant, width, well="gaussian", order=1.0, disorder=0.0, seed=0,
                    rotation=0.0, x0=0.0, y0=0.0, defects=()):
    """Sum of unit wells exp(-0.5 q), q = r^2/w^2 (q^order for super-Gaussian wells), depth
    1 + N(0, disorder) per site in site order when disorder > 0 (builtin_profiles.py:227-240)."""
    sites = lattice_sites(grid.lx, grid.ly, lattice, constant, rotation, x0, y0, defects)
    depth = (1.0 + np.random.default_rng(seed).normal(0.0, disorder, len(sites))) if disorder else np.ones(len(sites))
    out = np.zeros(grid.shape)
    for (_, _, _, x, y), d in zip(sites, depth):
        q = ((grid.X - x) ** 2 + (grid.Y - y) ** 2) / width ** 2
        out = out + d * np.exp(-0.5 * (q ** order if well == "supergaussian" else q))
    return out


def pulse_gaussian_coeff(t, P0=0.4, Pmax=1.2, sigma_time=0.1, cutoff_sigma=3.0,
                         pulse_separation=0.6, n_pulses=0, delay=0.0):
    """Peak-amplitude pulse-train coefficient (laser/pulse_gaussian.py:98-133)."""
    if t < delay:
        return 0.0
    t = t - delay
    phase = cutoff_sigma * sigma_time
    n = max(0, round((t - phase) / pulse_separation))
    if n_pulses > 0 and n >= n_pulses:
        return 0.0
    amp = min(P0 + n * (Pmax - P0), Pmax)
    d = t - n * pulse_separation - phase
    if abs(d) > cutoff_sigma * sigma_time:
        return 0.0
    return amp * math.exp(-0.5 * (d / sigma_time) ** 2)


# ---------------------------------------------------------------------------
# Physics constants (config/simulation_parameters.py:43-67 defaults)
# ---------------------------------------------------------------------------
@dataclass
class Physics:
    hbar: float = 0.6582119514
    m_eff: float = 0.284
    gamma_R: float = 0.005
    gamma_C: float = 0.083
    g_C: float = 0.001
    g_R: float = 0.002
    g_I: float = 0.0
    gamma_I: float = 0.001
    gamma_A: float = 0.005
    R: float = 0.02
    R_IA: float = 5e-2
    R_AI: float = 2.5e-3
    kappa: float = 0.05
    kinetic_relaxation_eta: float = 0.0
    alpha_2: float = 0.0
    zeeman: float = 0.0
    te_tm: float = 0.0

    def with_(self, **kw):
        return replace(self, **kw)


# ---------------------------------------------------------------------------
# Reservoirs
# ---------------------------------------------------------------------------
def reservoir_derivatives(kind, p, rho, pump, state):
    """Reservoir ODE right-hand sides with psi frozen.

    single            (single_reservoir.py:81-98):  dnR = P - (gR + R rho) nR
    double            (double_reservoir.py:255-279):
        dnI = P - (gI + R_IA) nI + R_AI nA ; dnA = R_IA nI - (gA + R_AI + R rho) nA
    quadratic-double  (quadratic_double_reservoir.py:131-187):
        dnR = kappa nI^2 - gR nR - R nR rho ; dnI = P - kappa nI^2 - gI nI
    """
    if kind == "none":
        return ()
    if kind == "single":
        (n,) = state
        return (pump - (p.gamma_R + p.R * rho) * n,)
    if kind == "double":
        nA, nI = state
        dI = pump - (p.gamma_I + p.R_IA) * nI + p.R_AI * nA
        dA = p.R_IA * nI - (p.gamma_A + p.R_AI + p.R * rho) * nA
        return (dA, dI)
    if kind == "quadratic-double":
        nR, nI = state
        transfer = p.kappa * nI * nI
        return (transfer - p.gamma_R * nR - p.R * nR * rho, pump - transfer - p.gamma_I * nI)
    raise ValueError(kind)


def active_inactive(kind, state):
    if kind == "none":
        return 0.0, 0.0
    if kind == "single":
        return state[0], 0.0
    return state[0], state[1]


def reservoir_rk2(kind, p, dt, psi, pump, state):
    """Midpoint RK2 with psi frozen, then clamp >= 0 (single_reservoir.py:100-111)."""
    if kind == "none":
        return ()
    rho = psi.real * psi.real + psi.imag * psi.imag
    k1 = reservoir_derivatives(kind, p, rho, pump, state)
    mid = tuple(s + 0.5 * dt * k for s, k in zip(state, k1))
    k2 = reservoir_derivatives(kind, p, rho, pump, mid)
    return tuple(np.maximum(s + dt * k, 0).astype(s.dtype, copy=False) for s, k in zip(state, k2))


# ---------------------------------------------------------------------------
# Finite-difference RHS (model.py:89-248)
# ---------------------------------------------------------------------------
def laplacian(f, dx, dy, boundary="periodic", kind="five-point"):
    """Finite-difference Laplacian; x is the last axis.

    five-point, periodic: model.py:214-225.  five-point, closed interval
    (mirrored Neumann edges written as 2 (f[1] - f[0]) / h^2): model.py:227-248.
    isotropic-9pt (dx == dy): rk4_cuda_solver.py:719-770, periodic wrap or the
    closed interval's mirrored neighbour indices.
    """
    if kind == "isotropic-9pt":
        return _laplacian9(f, dx, boundary)
    if boundary == "periodic":
        return ((np.roll(f, -1, axis=-2) - 2 * f + np.roll(f, 1, axis=-2)) / (dy ** 2)
                + (np.roll(f, -1, axis=-1) - 2 * f + np.roll(f, 1, axis=-1)) / (dx ** 2))
    if boundary != "closed-interval":
        raise ValueError(boundary)
    ny, nx = f.shape[-2:]
    out = np.zeros_like(f)
    c = (slice(None),) * (f.ndim - 2)
    if ny > 2 and nx > 2:
        out[c + (slice(1, -1), slice(1, -1))] = (
            (f[c + (slice(2, None), slice(1, -1))] - 2 * f[c + (slice(1, -1), slice(1, -1))]
             + f[c + (slice(None, -2), slice(1, -1))]) / dy ** 2
            + (f[c + (slice(1, -1), slice(2, None))] - 2 * f[c + (slice(1, -1), slice(1, -1))]
               + f[c + (slice(1, -1), slice(None, -2))]) / dx ** 2)
    if nx > 2:
        for row, nb in ((0, 1), (-1, -2)):
            out[c + (row, slice(1, -1))] = (
                2 * (f[c + (nb, slice(1, -1))] - f[c + (row, slice(1, -1))]) / dy ** 2
                + (f[c + (row, slice(2, None))] - 2 * f[c + (row, slice(1, -1))]
                   + f[c + (row, slice(None, -2))]) / dx ** 2)
    if ny > 2:
        for col, nb in ((0, 1), (-1, -2)):
            out[c + (slice(1, -1), col)] = (
                (f[c + (slice(2, None), col)] - 2 * f[c + (slice(1, -1), col)]
                 + f[c + (slice(None, -2), col)]) / dy ** 2
                + 2 * (f[c + (slice(1, -1), nb)] - f[c + (slice(1, -1), col)]) / dx ** 2)
    for row, nby in ((0, 1), (-1, -2)):
        for col, nbx in ((0, 1), (-1, -2)):
            out[c + (row, col)] = (2 * (f[c + (nby, col)] - f[c + (row, col)]) / dy ** 2
                                   + 2 * (f[c + (row, nbx)] - f[c + (row, col)]) / dx ** 2)
    return out


def _neighbour(n, boundary):
    """Index arrays of the -1 / +1 neighbours along an axis of length n."""
    i = np.arange(n)
    if boundary == "periodic":
        return (i - 1) % n, (i + 1) % n
    lo = np.where(i > 0, i - 1, 1 if n > 1 else 0)
    hi = np.where(i < n - 1, i + 1, n - 2 if n > 1 else 0)
    return lo, hi


def _laplacian9(f, h, boundary):
    """(4 (dn + up + rt + lt) + ul + ur + dl + dr - 20 f) / (6 h^2) in the kernel's order."""
    ny, nx = f.shape[-2:]
    up, dn = _neighbour(ny, boundary)
    lt, rt = _neighbour(nx, boundary)
    rows = lambda idx: np.take(f, idx, axis=-2)
    cols = lambda a, idx: np.take(a, idx, axis=-1)
    fdn, fup = rows(dn), rows(up)
    inv_6h2 = (1.0 / (h * h)) / 6.0
    card = ((fdn + fup) + cols(f, rt)) + cols(f, lt)
    acc = 4.0 * card + cols(fup, lt)
    acc = acc + cols(fup, rt)
    acc = acc + cols(fdn, lt)
    acc = acc + cols(fdn, rt)
    return (acc - 20.0 * f) * inv_6h2


def te_tm_energy(p, field, dx, dy, boundary="periodic"):
    """TE-TM term -beta[(dxx - dyy) -/+ 2i dxy] psi_-/+ (model.py:175-208).

    The second differences act on the other component: periodic wrap (model.py:187-194), or on
    the closed interval the field extended by one mirrored point per side, f[-1] = f[1],
    f[n] = f[n - 2], corners mirrored in both axes (model.py:196-205, numpy's "reflect" pad).
    """
    if p.te_tm == 0.0:
        return 0
    other = field[:, ::-1]
    if boundary == "periodic":
        dxx = (np.roll(other, -1, axis=-1) - 2 * other + np.roll(other, 1, axis=-1)) / dx ** 2
        dyy = (np.roll(other, -1, axis=-2) - 2 * other + np.roll(other, 1, axis=-2)) / dy ** 2
        dxy = (np.roll(np.roll(other, -1, axis=-2), -1, axis=-1)
               - np.roll(np.roll(other, -1, axis=-2), 1, axis=-1)
               - np.roll(np.roll(other, 1, axis=-2), -1, axis=-1)
               + np.roll(np.roll(other, 1, axis=-2), 1, axis=-1)) / (4 * dx * dy)
    elif boundary == "closed-interval":
        ny, nx = other.shape[-2:]
        up, dn = _neighbour(ny, boundary)   # rows y - 1, y + 1 (mirrored at the walls)
        lt, rt = _neighbour(nx, boundary)
        rows = lambda idx: np.take(other, idx, axis=-2)
        cols = lambda a, idx: np.take(a, idx, axis=-1)
        dxx = (cols(other, rt) - 2 * other + cols(other, lt)) / dx ** 2
        dyy = (rows(dn) - 2 * other + rows(up)) / dy ** 2
        dxy = (cols(rows(dn), rt) - cols(rows(dn), lt) - cols(rows(up), rt) + cols(rows(up), lt)) / (4 * dx * dy)
    else:
        raise ValueError(boundary)
    cross = np.stack((-2j * dxy[:, 0], 2j * dxy[:, 1]), axis=1)
    return -p.te_tm * (dxx - dyy + cross)


def fdm_rhs(kind, p, grid, psi, res, pump, potential=0.0, pump_weights=None,
            reservoir_components=1, laplacian_kind="five-point"):
    """Scalar or spinor FDM right-hand side; returns (dpsi, dres).

    Scalar: model.py:89-127.  Spinor: model.py:129-173 (Zeeman +/-Z/2,
    alpha_2 cross density, TE-TM 9-point stencil).
    """
    spinor = psi.shape[1] == 2
    lap = laplacian(psi, grid.dx, grid.dy, getattr(grid, "boundary", "periodic"), laplacian_kind)
    if spinor:
        if reservoir_components == 1:
            rho_res = (psi.real * psi.real + psi.imag * psi.imag).sum(axis=1, keepdims=True)
            dres = reservoir_derivatives(kind, p, rho_res, pump, res)
        else:
            rho_res = psi.real * psi.real + psi.imag * psi.imag
            dres = reservoir_derivatives(kind, p, rho_res, pump * pump_weights, res)
    else:
        rho_res = psi.real * psi.real + psi.imag * psi.imag
        dres = reservoir_derivatives(kind, p, rho_res, pump, res)
    nA, nI = active_inactive(kind, res)
    density = np.abs(psi) ** 2
    kinetic = -(p.hbar ** 2) / (2 * p.m_eff) * lap
    eff = potential + p.g_C * density + p.g_R * nA + p.g_I * nI
    extra = 0
    if spinor:
        sign = np.asarray([1, -1], dtype=density.dtype).reshape(1, 2, 1, 1)
        eff = eff + p.alpha_2 * density[:, ::-1] + p.zeeman * sign / 2
        extra = te_tm_energy(p, psi, grid.dx, grid.dy, getattr(grid, "boundary", "periodic"))
    gain = (p.R * nA - p.gamma_C) / 2.0
    hb2m = p.hbar / (2.0 * p.m_eff)
    dpsi = ((-1j / p.hbar) * (kinetic + eff * psi + extra) + gain * psi
            + p.kinetic_relaxation_eta * nA * hb2m * lap)
    return dpsi, dres


def rk4_fdm_step(kind, p, grid, psi, res, pump_at, t, dt, **kw):
    """Classic stage-coupled RK4 (rk4_fdm_solver.py:36-77); reservoir clamped >= 0."""
    k1, r1 = fdm_rhs(kind, p, grid, psi, res, pump_at(t), **kw)
    k2, r2 = fdm_rhs(kind, p, grid, psi + 0.5 * dt * k1,
                     tuple(a + 0.5 * dt * b for a, b in zip(res, r1)), pump_at(t + 0.5 * dt), **kw)
    k3, r3 = fdm_rhs(kind, p, grid, psi + 0.5 * dt * k2,
                     tuple(a + 0.5 * dt * b for a, b in zip(res, r2)), pump_at(t + 0.5 * dt), **kw)
    k4, r4 = fdm_rhs(kind, p, grid, psi + dt * k3,
                     tuple(a + dt * b for a, b in zip(res, r3)), pump_at(t + dt), **kw)
    psi_new = (psi + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)).astype(psi.dtype, copy=False)
    res_new = tuple(
        np.maximum(r0 + (dt / 6.0) * (a + 2 * b + 2 * c + d), 0).astype(r0.dtype, copy=False)
        for r0, a, b, c, d in zip(res, r1, r2, r3, r4))
    return psi_new, res_new


# ---------------------------------------------------------------------------
# Split-step FFT
# ---------------------------------------------------------------------------
def _nonlinear_exact(p, psi, res, kind, dt, potential=0.0):
    """Exact local gain/phase step (split_step_fft_solver.py:124-140)."""
    nA, nI = active_inactive(kind, res)
    rho = np.abs(psi) ** 2
    gain = (p.R * nA - p.gamma_C) / 2.0 + np.imag(potential) / p.hbar
    x = 2 * gain * dt
    x = np.asarray(x, dtype=np.float64) if np.ndim(x) == 0 else x
    small = np.abs(x) < 1e-6
    phase_time = dt * np.where(small, 1 + x / 2 + x * x / 6, np.expm1(x) / np.where(small, 1, x))
    phase = ((np.real(potential) + p.g_R * nA + p.g_I * nI) * dt + p.g_C * rho * phase_time) / p.hbar
    return psi * np.exp(gain * dt - 1j * phase)


def split_step_cpu(kind, p, grid, psi, res, pump_at, t, dt, step, noise=None):
    """Strang step R(dt/2) K(dt/2) N(dt) noise K(dt/2) R(dt/2) (split_step_fft_solver.py:97-159).

    ``noise`` is a dict for :func:`wigner_add` (seeds, gamma, scattering,
    strength, volume, precision) or None; it uses nR after the first half.
    """
    cdt = psi.dtype
    half = np.exp(-1j * p.hbar * grid.k_squared * (dt / 2) / (2 * p.m_eff))
    res = reservoir_rk2(kind, p, dt / 2, psi, pump_at(t + dt / 4), res)
    psi = np.fft.ifft2(np.fft.fft2(psi) * half).astype(cdt)
    psi = _nonlinear_exact(p, psi, res, kind, dt).astype(cdt)
    if noise is not None:
        active = res[0] if res else 0
        wigner_add(psi, active, dt=dt, step=step, **noise)
    psi = np.fft.ifft2(np.fft.fft2(psi) * half).astype(cdt)
    res = reservoir_rk2(kind, p, dt / 2, psi, pump_at(t + 3 * dt / 4), res)
    return psi, res


def split_step_kouter(kind, p, grid, psi, res, pump_at, t, dt, step, noise=None):
    """K(dt/2) [R(dt/2) N(dt) R(dt/2) + noise] K(dt/2) (tests/crosscheck/test_split_step_fft_cuda.py:46-70;
    split_step_fft_cuda_solver.py:106-184 uses nR after the second half for noise)."""
    cdt = psi.dtype
    half = np.exp(-1j * p.hbar * grid.k_squared * dt / (4 * p.m_eff))
    psi = np.fft.ifft2(np.fft.fft2(psi) * half).astype(cdt)
    res = reservoir_rk2(kind, p, dt / 2, psi, pump_at(t + dt / 4), res)
    psi = _nonlinear_exact(p, psi, res, kind, dt).astype(cdt)
    res = reservoir_rk2(kind, p, dt / 2, psi, pump_at(t + 3 * dt / 4), res)
    if noise is not None:
        active = res[0] if res else 0
        wigner_add(psi, active, dt=dt, step=step, **noise)
    psi = np.fft.ifft2(np.fft.fft2(psi) * half).astype(cdt)
    return psi, res


def _spinor_reservoir_half(kind, p, dt, psi, pump, res, weights, res_components):
    """Spinor reservoir half-step (split_step_fft_solver.py:81-94): one reservoir per component
    driven by |psi_c|^2 and P w_c, or one shared reservoir driven by |psi+|^2 + |psi-|^2 and P;
    midpoint rule with psi frozen, then clamp >= 0."""
    if kind == "none":
        return ()
    if res_components == 2:
        return reservoir_rk2(kind, p, dt, psi, pump * weights, res)
    rho = (psi.real * psi.real + psi.imag * psi.imag).sum(axis=1, keepdims=True)
    k1 = reservoir_derivatives(kind, p, rho, pump, res)
    mid = tuple(s + 0.5 * dt * k for s, k in zip(res, k1))
    k2 = reservoir_derivatives(kind, p, rho, pump, mid)
    return tuple(np.maximum(s + dt * k, 0).astype(s.dtype, copy=False) for s, k in zip(res, k2))


def _spinor_nonlinear(p, psi, res, kind, dt):
    """Exact local step of both components (split_step_fft_solver.py:124-143): the scalar step
    per component plus the cross-spin phase alpha_2 |psi_o|^2 tau_o / hbar of the other one."""
    nA, nI = active_inactive(kind, res)
    rho = np.abs(psi) ** 2
    gain = (p.R * nA - p.gamma_C) / 2.0
    x = 2 * gain * dt
    x = np.asarray(x, dtype=np.float64) if np.ndim(x) == 0 else x
    small = np.abs(x) < 1e-6
    phase_time = dt * np.where(small, 1 + x / 2 + x * x / 6, np.expm1(x) / np.where(small, 1, x))
    phase = ((p.g_R * nA + p.g_I * nI) * dt + p.g_C * rho * phase_time) / p.hbar
    other = np.broadcast_to(phase_time, rho.shape)[:, ::-1]
    phase = phase + p.alpha_2 * rho[:, ::-1] * other / p.hbar
    return psi * np.exp(gain * dt - 1j * phase)


def _spinor_kinetic_half(grid, p, psi, dt):
    """FFT, exp(-i H(k) dt / (2 hbar)) as a complex matrix of the field precision, IFFT
    (split_step_fft_solver.py:46-53, 71-79)."""
    cdt = psi.dtype
    M = spinor_propagator(grid, p, dt / 2).astype(cdt)
    psi_k = np.fft.fft2(psi, axes=(-2, -1)).astype(cdt)
    return np.fft.ifft2(apply_2x2(M, psi_k), axes=(-2, -1)).astype(cdt)


def split_step_spinor(kind, p, grid, psi, res, pump_at, t, dt, step, weights, res_components, noise=None,
                      kouter=False):
    """Spinor split step, (1, 2, ny, nx) psi.  CPU order R K N noise K R
    (split_step_fft_solver.py:97-159); ``kouter``: K [R N R noise] K, the order of the
    reference's spinor fast path (split_step_fft_cuda_solver.py:229-330)."""
    cdt = psi.dtype
    half = lambda r, pump: _spinor_reservoir_half(kind, p, dt / 2, psi, pump, r, weights, res_components)
    if not kouter:
        res = half(res, pump_at(t + dt / 4))
        psi = _spinor_kinetic_half(grid, p, psi, dt)
        psi = _spinor_nonlinear(p, psi, res, kind, dt).astype(cdt)
        if noise is not None:
            wigner_add(psi, res[0] if res else 0, dt=dt, step=step, **noise)
        psi = _spinor_kinetic_half(grid, p, psi, dt)
        res = half(res, pump_at(t + 3 * dt / 4))
        return psi, res
    psi = _spinor_kinetic_half(grid, p, psi, dt)
    res = half(res, pump_at(t + dt / 4))
    psi = _spinor_nonlinear(p, psi, res, kind, dt).astype(cdt)
    res = _spinor_reservoir_half(kind, p, dt / 2, psi, pump_at(t + 3 * dt / 4), res, weights, res_components)
    if noise is not None:
        wigner_add(psi, res[0] if res else 0, dt=dt, step=step, **noise)
    psi = _spinor_kinetic_half(grid, p, psi, dt)
    return psi, res


# ---------------------------------------------------------------------------
# Tableaus (independent source: SciPy's published Dormand-Prince / DOP853 data)
# ---------------------------------------------------------------------------
@dataclass
class OracleTableau:
    name: str
    order: int
    embedded_order: int | None
    A: np.ndarray
    b: np.ndarray
    c: np.ndarray
    b_hat: np.ndarray | None = None
    b_hat3: np.ndarray | None = None


def oracle_tableau(name: str) -> OracleTableau:
    """RK4 by definition; DP5 and DOP853 from scipy.integrate's coefficient tables."""
    if name == "rk4":
        A = np.zeros((4, 4))
        A[1, 0] = 0.5
        A[2, 1] = 0.5
        A[3, 2] = 1.0
        return OracleTableau("rk4", 4, None, A, np.array([1, 2, 2, 1]) / 6.0,
                             np.array([0, 0.5, 0.5, 1.0]))
    from scipy.integrate._ivp import rk as _rk
    if name == "dp5":
        s = _rk.RK45
        A = np.zeros((7, 7))
        A[:6, :5] = s.A
        A[6, :6] = s.B
        b = np.append(s.B, 0.0)
        c = np.append(s.C, 1.0)
        return OracleTableau("dp5", 5, 4, A, b, c, b_hat=b - s.E)
    if name == "dop853":
        from scipy.integrate._ivp import dop853_coefficients as d
        n = d.N_STAGES
        A = np.array(d.A[:n, :n], dtype=np.float64)
        b = np.array(d.B, dtype=np.float64)
        c = np.array(d.C[:n], dtype=np.float64)
        return OracleTableau("dop853", 8, 5, A, b, c, b_hat=b - d.E5[:n], b_hat3=b - d.E3[:n])
    raise ValueError(name)


# ---------------------------------------------------------------------------
# Lawson IFRK (solver/ifrk_fft_cuda_solver.py:96-293)
# ---------------------------------------------------------------------------
@dataclass
class LawsonProblem:
    kind: str
    p: Physics
    grid: Grid
    precision: str = "double"
    potential: object = 0.0
    tableau: str = "rk4"
    adaptive: bool = False
    rtol: float = 1e-6
    atol: float = 1e-8
    pump_weights: object = None
    reservoir_components: int = 1
    _tab: OracleTableau = field(init=False, default=None)

    def __post_init__(self):
        self._tab = oracle_tableau(self.tableau)
        self.real = np.float32 if self.precision == "single" else np.float64
        self.cplx = np.complex64 if self.precision == "single" else np.complex128
        self.k2 = np.asarray(self.grid.k_squared, dtype=self.real)

    # dispersion.py:34-83 for spinors; scalar phase exp(i*(-hbar/2m)*h*f*k^2)
    def linear(self, fraction, h, fk):
        p = self.p
        if fk.shape[1] == 2:
            tau = float(h * fraction)
            cache = self.__dict__.setdefault("_spinor_cache", {})
            M = cache.get(tau)
            if M is None:
                if len(cache) > 48:
                    cache.clear()
                M = cache[tau] = spinor_propagator(self.grid, p, tau).astype(self.cplx)
            return apply_2x2(M, fk)
        rate = self.real(-p.hbar / (2.0 * p.m_eff))
        angle = rate * self.real(h * fraction) * self.k2
        return (np.exp(1j * angle).astype(self.cplx, copy=False) * fk)

    def rhs_psi(self, psi, res, rho, psi_k):
        p = self.p
        nA, nI = active_inactive(self.kind, res)
        eff = self.potential + p.g_C * rho + p.g_R * nA + p.g_I * nI
        if psi.shape[1] == 2:
            eff = eff + p.alpha_2 * rho[:, ::-1]
        gain = 0.5 * (p.R * nA - p.gamma_C)
        out = (-1j / p.hbar) * eff * psi + gain * psi
        if p.kinetic_relaxation_eta != 0.0:
            lap = np.fft.ifft2(-self.k2 * psi_k, axes=(-2, -1))
            out = out + p.kinetic_relaxation_eta * nA * (p.hbar / (2.0 * p.m_eff)) * lap
        return out.astype(self.cplx, copy=False)

    def res_derivs(self, psi, pump, res, rho):
        if psi.shape[1] == 2:
            if self.reservoir_components == 2:
                return reservoir_derivatives(self.kind, self.p, rho, pump * self.pump_weights, res)
            return reservoir_derivatives(self.kind, self.p, rho.sum(axis=1, keepdims=True), pump, res)
        return reservoir_derivatives(self.kind, self.p, rho, pump, res)

    def step(self, psi0, res0, pump_at, t, h, noise=None, step=0):
        """One Lawson trial.  Returns (accepted, error, psi_new, res_new).

        ``noise`` (a :func:`wigner_add` keyword dict without dt / step) adds the truncated-Wigner
        increment of step index ``step`` to the combined field in real space, with the final
        reservoir as the active density, and returns to k-space from there
        (solver/ifrk_cuda_fused.py:908-923: inverse transform, add_wigner on res_out[0], forward
        transform).  The fused fixed-step path is the only one that injects it."""
        tab = self._tab
        sc = self.real
        fft = lambda a: np.fft.fft2(a, axes=(-2, -1)).astype(self.cplx, copy=False)
        ifft = lambda a: np.fft.ifft2(a, axes=(-2, -1)).astype(self.cplx, copy=False)
        psi0_k = fft(psi0)
        f, derivs, pumps = [], [], {}
        for i, ci in enumerate(tab.c):
            ci = float(ci)
            if i == 0:
                psi_i, psi_i_k = psi0, psi0_k
            else:
                psi_i_k = self.linear(ci, h, psi0_k)
                for j in range(i):
                    aij = float(tab.A[i, j])
                    if aij:
                        psi_i_k = psi_i_k + sc(h * aij) * self.linear(ci - float(tab.c[j]), h, f[j])
                psi_i_k = psi_i_k.astype(self.cplx, copy=False)
                psi_i = ifft(psi_i_k)
            res_i = []
            for fi, r0 in enumerate(res0):
                v = r0.copy()
                for j in range(i):
                    aij = float(tab.A[i, j])
                    if aij:
                        v += sc(h * aij) * derivs[j][fi]
                res_i.append(v.astype(self.real, copy=False))
            res_i = tuple(res_i)
            rho = psi_i.real * psi_i.real + psi_i.imag * psi_i.imag
            if ci not in pumps:
                pumps[ci] = pump_at(t + ci * h)
            f.append(fft(self.rhs_psi(psi_i, res_i, rho, psi_i_k)))
            derivs.append(self.res_derivs(psi_i, pumps[ci], res_i, rho))
        psi_new_k = self.linear(1.0, h, psi0_k)
        for j, bj in enumerate(tab.b):
            bj = float(bj)
            if bj:
                psi_new_k = psi_new_k + sc(h * bj) * self.linear(1.0 - float(tab.c[j]), h, f[j])
        psi_new_k = psi_new_k.astype(self.cplx, copy=False)
        res_new = []
        for fi, r0 in enumerate(res0):
            v = r0.copy()
            for j, bj in enumerate(tab.b):
                if float(bj):
                    v += sc(h * float(bj)) * derivs[j][fi]
            res_new.append(np.maximum(v, 0).astype(self.real, copy=False))
        res_new = tuple(res_new)
        psi_new = ifft(psi_new_k)
        if noise is not None:
            wigner_add(psi_new, res_new[0] if res_new else 0, dt=h, step=step, **noise)
        if not self.adaptive:
            return True, None, psi_new, res_new
        error = self.error(h, psi0, psi_new, psi0_k, f, res0, res_new, derivs)
        return error <= 1, error, psi_new, res_new

    def error(self, h, psi_start, psi_new, psi0_k, f, res0, res_new, derivs):
        """Embedded error norm, max over members (ifrk_fft_cuda_solver.py:243-293)."""
        tab = self._tab
        sc = self.real
        axes = tuple(range(1, psi_new.ndim))

        def spectral(coeffs):
            acc = np.zeros_like(psi0_k)
            for j, cj in enumerate(coeffs):
                if cj:
                    acc = acc + sc(h * cj) * self.linear(1.0 - float(tab.c[j]), h, f[j])
            return acc

        def estimator(coeffs, ferr):
            scale = self.atol + self.rtol * np.maximum(np.abs(psi_start), np.abs(psi_new))
            mag = np.abs(ferr)
            s = np.where(scale > 0, mag / np.where(scale > 0, scale, 1), np.where(mag == 0, 0, np.inf))
            sq = np.sum(s * s, axis=axes)
            count = psi_new[0].size
            for fi, (r0, r1) in enumerate(zip(res0, res_new)):
                v = np.zeros_like(r0)
                for j, cj in enumerate(coeffs):
                    if cj:
                        v = v + sc(h * cj) * derivs[j][fi]
                scale = self.atol + self.rtol * np.maximum(np.abs(r0), np.abs(r1))
                mag = np.abs(v)
                s = np.where(scale > 0, mag / np.where(scale > 0, scale, 1), np.where(mag == 0, 0, np.inf))
                sq = sq + np.sum(s * s, axis=axes)
                count += r0[0].size
            return sq, count

        w5 = [float(b) - float(bh) for b, bh in zip(tab.b, tab.b_hat)]
        e5field = np.fft.ifft2(spectral(w5), axes=(-2, -1))
        err5, count = estimator(w5, e5field)
        if tab.b_hat3 is not None:
            w3 = [float(b) - float(bh) for b, bh in zip(tab.b, tab.b_hat3)]
            e3field = np.fft.ifft2(spectral(w3), axes=(-2, -1))
            err3, _ = estimator(w3, e3field)
            den = err5 + 0.01 * err3
            member = np.where(den > 0, err5 / np.sqrt(np.where(den > 0, den, 1) * count), 0)
        else:
            member = np.sqrt(err5 / count)
        return float(np.max(member))


def ip_rk4_step(prob: LawsonProblem, psi0, res0, pump_at, t, dt):
    """One interaction-picture RK4 step (solver/ip_rk4_solver.py:122-213): the state is carried
    as u = exp(-L s) psi^k; stages at the midpoint use exp(+-L dt/2), the end exp(+-L dt); the
    propagators are formed in complex128 and rounded to the field precision
    (ip_rk4_solver.py:45-71).  Reservoir stages are coupled as in the Lawson step."""
    p, g, cplx = prob.p, prob.grid, prob.cplx
    fft = lambda a: np.fft.fft2(a, axes=(-2, -1)).astype(cplx, copy=False)
    ifft = lambda a: np.fft.ifft2(a, axes=(-2, -1)).astype(cplx, copy=False)
    spinor = psi0.shape[1] == 2
    if spinor:
        E = {s: spinor_propagator(g, p, s * dt).astype(cplx) for s in (0.5, 1.0, -0.5, -1.0)}
        lin = lambda s, x: apply_2x2(E[s], x).astype(cplx, copy=False)
    else:
        L = -1j * p.hbar * g.k_squared / (2 * p.m_eff)
        E = {s: np.exp(L * (s * dt)).astype(cplx) for s in (0.5, 1.0, -0.5, -1.0)}
        lin = lambda s, x: (E[s] * x).astype(cplx, copy=False)
    rho_of = lambda a: a.real * a.real + a.imag * a.imag
    p0, p_half, p1 = pump_at(t), pump_at(t + 0.5 * dt), pump_at(t + dt)

    def stage(psi, psi_k, res, pump):
        rho = rho_of(psi)
        return fft(prob.rhs_psi(psi, res, rho, psi_k)), prob.res_derivs(psi, pump, res, rho)

    shift = lambda res, ks, h: tuple((r + h * k).astype(prob.real, copy=False) for r, k in zip(res, ks))
    psi0_k = fft(psi0)
    k1, kr1 = stage(psi0, psi0_k, res0, p0)
    mid_k = lin(0.5, psi0_k + (dt / 2) * k1)
    res_mid = shift(res0, kr1, 0.5 * dt)
    n2, kr2 = stage(ifft(mid_k), mid_k, res_mid, p_half)
    k2 = lin(-0.5, n2)
    mid2_k = lin(0.5, psi0_k + (dt / 2) * k2)
    res_mid2 = shift(res0, kr2, 0.5 * dt)
    n3, kr3 = stage(ifft(mid2_k), mid2_k, res_mid2, p_half)
    k3 = lin(-0.5, n3)
    end_k = lin(1.0, psi0_k + dt * k3)
    res_end = shift(res0, kr3, dt)
    n4, kr4 = stage(ifft(end_k), end_k, res_end, p1)
    k4 = lin(-1.0, n4)
    psi_new = ifft(lin(1.0, psi0_k + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)))
    res_new = tuple(np.maximum(r + (dt / 6.0) * (a + 2 * b + 2 * c + d), 0).astype(prob.real, copy=False)
                    for r, a, b, c, d in zip(res0, kr1, kr2, kr3, kr4))
    return psi_new, res_new


class StepController:
    """PI controller restated from solver/step_controller.py:8-81."""

    def __init__(self, dt, tableau_order, embedded_order, dt_max=None):
        self.dt_max = dt_max
        self.h_next = min(dt, dt_max or math.inf)
        self.k = min(tableau_order, embedded_order) + 1
        self.err_prev = 1.0
        self.no_growth = False
        self.accepted = 0
        self.rejected = 0

    def propose(self, t, target):
        prop = self.h_next
        clamped = prop >= target - t
        h = target - t if clamped else prop
        if h < 16 * np.finfo(float).eps * max(1, abs(t)):
            raise RuntimeError(f"adaptive step size underflow at t={t}")
        return h, clamped, prop

    def update(self, accepted, error, h, clamped, prop):
        if error is None or not math.isfinite(error) or error < 0:
            error = math.inf
        if accepted:
            self.accepted += 1
            factor = 5.0 if error == 0 else 0.9 * error ** (-0.7 / self.k) * self.err_prev ** (0.4 / self.k)
            factor = min(5.0, max(0.2, factor))
            if self.no_growth:
                factor = min(1.0, factor)
            self.err_prev = max(error, 1e-4)
            self.no_growth = False
        else:
            self.rejected += 1
            factor = min(1.0, max(0.2, 0.9 * error ** (-1 / self.k)))
            self.no_growth = True
        h_next = h * factor
        if accepted and clamped:
            h_next = max(h_next, prop)
        self.h_next = min(h_next, self.dt_max or math.inf)


def lawson_advance_adaptive(prob: LawsonProblem, psi, res, pump_at, t0, dt, n_ticks, ctl=None, chunk=1):
    """Adaptive trials in advances of ``chunk`` ticks (abstract_solver.py:86-96).

    Each advance runs trials until ``t`` reaches its target ``t0 + (step + chunk) dt``;
    only that target is landed on exactly (the controller clamps the last trial).
    The reference's run loop advances one tick at a time when it records every step.
    """
    tab = prob._tab
    if ctl is None:
        ctl = StepController(dt, tab.order, tab.embedded_order)
    t = t0
    step = 0
    while step < n_ticks:
        n = min(chunk, n_ticks - step)
        target = t0 + (step + n) * dt
        while t < target:
            h, clamped, prop = ctl.propose(t, target)
            ok, err, p_new, r_new = prob.step(psi, res, pump_at, t, h)
            if ok:
                psi, res = p_new, r_new
                t = target if clamped else t + h
            ctl.update(ok, err, h, clamped, prop)
        step += n
        t = t0 + step * dt
    return psi, res, ctl


# ---------------------------------------------------------------------------
# Spinor dispersion (dispersion.py:34-89)
# ---------------------------------------------------------------------------
def spinor_hamiltonian(grid, p):
    """H(k) = [[E + Z/2, beta (kx - i ky)^2], [beta (kx + i ky)^2, E - Z/2]]."""
    E = p.hbar ** 2 * grid.k_squared / (2 * p.m_eff)
    coupling = p.te_tm * (grid.KX - 1j * grid.KY) ** 2
    H = np.empty((2, 2) + grid.shape, dtype=np.complex128)
    H[0, 0] = E + p.zeeman / 2
    H[1, 1] = E - p.zeeman / 2
    H[0, 1] = coupling
    H[1, 0] = np.conj(coupling)
    return H


def spinor_propagator(grid, p, tau_t):
    """exp(-i H t / hbar) by the traceless closed form with sinc (dispersion.py:66-83)."""
    H = spinor_hamiltonian(grid, p)
    eps = (H[0, 0] + H[1, 1]) / 2
    z = (H[0, 0] - H[1, 1]) / 2
    omega = np.sqrt(z * z + H[0, 1] * H[1, 0])
    tau = tau_t / p.hbar
    theta = omega * tau
    ph = np.exp(-1j * eps * tau)
    cs = np.cos(theta)
    so = tau * np.sinc(theta / np.pi)
    M = np.empty_like(H)
    M[0, 0] = ph * (cs - 1j * so * z)
    M[1, 1] = ph * (cs + 1j * so * z)
    M[0, 1] = -1j * ph * so * H[0, 1]
    M[1, 0] = -1j * ph * so * H[1, 0]
    return M


def apply_2x2(M, f):
    out = np.empty_like(f)
    out[:, 0] = M[0, 0] * f[:, 0] + M[0, 1] * f[:, 1]
    out[:, 1] = M[1, 0] * f[:, 0] + M[1, 1] * f[:, 1]
    return out


# ---------------------------------------------------------------------------
# Steady state (simulation_controller.py:710-734)
# ---------------------------------------------------------------------------
def steady_metrics(psi, old_density, cell):
    """Return (density, norms, relative density change, relative norm change)."""
    density = np.abs(psi) ** 2
    if psi.shape[1] == 2:
        density = density.sum(axis=1)
    else:
        density = density[:, 0]
    axes = (-2, -1)
    norms = density.sum(axis=axes) * cell
    if old_density is None:
        return density, norms, None, None
    delta = np.sqrt(((density - old_density) ** 2).sum(axis=axes))
    scale = np.sqrt((old_density ** 2).sum(axis=axes))
    old_norms = old_density.sum(axis=axes) * cell
    rel_d = delta / np.maximum(scale, 1e-30)
    rel_n = np.abs(norms - old_norms) / np.maximum(old_norms, 1e-30)
    return density, norms, rel_d, rel_n


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def threshold_single(p: Physics) -> float:
    """Pump threshold of the single reservoir: P_th = gamma_R gamma_C / R."""
    return p.gamma_R * p.gamma_C / p.R


def threshold_double(p: Physics) -> float:
    """Threshold of the double reservoir at psi = 0: nA_ss = gamma_C / R with
    nA_ss = R_IA P / [(gA + R_AI)(gI + R_IA) - R_IA R_AI] (double_reservoir.py:268-269)."""
    return (p.gamma_C / p.R) * ((p.gamma_A + p.R_AI) * (p.gamma_I + p.R_IA) / p.R_IA - p.R_AI)
