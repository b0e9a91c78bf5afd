"""Training data, training loop and warm starts of the FNO surrogate.

* ``make_members``  -- seeded sweep members (256 x 256, lx = ly = 128 um,
  single reservoir): random pumps from ``polarism_amd.pumps`` with peak
  P0 = U[0.8, 4] x the uniform-pump threshold.
* ``generate``      -- runs the members from noise-seeded cold starts with
  ``PackedSweep`` (rk4-fdm-hip, complex64) to a steady state and returns the
  surrogate's inputs and targets.  Targets are gauge fixed: the global phase
  is chosen so that sum psi |psi|^2 is real and positive (the steady state
  of the U(1)-symmetric GPE is defined only up to a global phase).
* ``train``         -- AdamW on the mean squared error of normalised targets;
  the network's forward and backward run on the HIP kernels.
* ``warm_start``    -- predicted (psi, n_R) as the initial state of the same
  steady-state run; a member predicted to hold no condensate (peak density
  below ``ZERO_STATE_PEAK``) starts from psi = 0 with the predicted reservoir.
  psi = 0 is a fixed point of the noiseless dynamics, so such a member meets the
  criterion at once and a wrong "no condensate" prediction cannot correct itself:
  the noise-seeded cold start of the same member is the judge of every warm
  result (a member counts only where both converge to the same density).

The inputs carry no coordinate channels, so the network is equivariant to
the translations of the periodic box, and training augments every batch with
a random circular shift and one of the eight symmetries of the square (the
GPE on a square periodic grid with dx = dy is invariant under both).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from ..config import PhysicsConstants
from ..pumps import random_pump, threshold
from ..sweep import PackedSweep, SweepMember

GRID = dict(n=256, L=128.0)
IN_FEATURES = 4
# A member whose predicted peak density stays below this (|psi|^2, in the units of the solver)
# is started from psi = 0, the steady state of a pump below threshold: a predicted residue of
# order 1e-3 would otherwise have to decay to the norm < 1e-30 test like a cold start does.
# (Starting such a member from the cold start's noise field instead was measured slower than
# the cold start itself: with the predicted reservoir already at P / gamma_R the noise decays
# more slowly than from an empty reservoir.)  Condensed steady states of the training
# distribution peak above ~0.05; the surrogate's residue on below-threshold members stays
# under ~0.01.
ZERO_STATE_PEAK = 1e-2


def make_members(count, seed, n=GRID["n"], L=GRID["L"], ratio=(0.8, 4.0)):
    p = PhysicsConstants()
    rng = np.random.default_rng(seed)
    pth = threshold(p, "single")
    out = []
    for i in range(count):
        env, desc = random_pump(rng, n, n, L, L)
        r = float(rng.uniform(*ratio))
        out.append(SweepMember(n, n, L, L, p, env, r * pth, seed=seed * 100003 + i,
                               label={"ratio": r, "kind": desc["kind"]}))
    return out


def features(members) -> np.ndarray:
    """(4, B, H, W) float32: local pump / P_th, P0 / P_th, gamma_C / 0.083, g_C / 0.001."""
    m0 = members[0]
    H, W = m0.ny, m0.nx
    out = np.empty((IN_FEATURES, len(members), H, W), dtype=np.float32)
    for b, m in enumerate(members):
        pth = threshold(m.physics, "single")
        ratio = m.P0 / pth
        out[0, b] = m.envelope * ratio
        out[1, b] = ratio
        out[2, b] = m.physics.gamma_C / 0.083
        out[3, b] = m.physics.g_C / 0.001
    return out


def gauge_fix(psi: np.ndarray) -> np.ndarray:
    w = np.sum(psi * np.abs(psi) ** 2)
    if abs(w) == 0:
        return psi
    return psi * np.exp(-1j * np.angle(w))


@dataclass
class Normaliser:
    psi_scale: float
    n_scale: float

    def targets(self, psi_list, n_list) -> np.ndarray:
        B = len(psi_list)
        H, W = psi_list[0].shape
        out = np.empty((3, B, H, W), dtype=np.float32)
        for b in range(B):
            z = gauge_fix(psi_list[b].astype(np.complex128))
            out[0, b] = z.real / self.psi_scale
            out[1, b] = z.imag / self.psi_scale
            out[2, b] = n_list[b] / self.n_scale
        return out

    def fields(self, y: np.ndarray, b: int):
        psi = (y[0, b] + 1j * y[1, b]) * self.psi_scale
        n = np.maximum(y[2, b] * self.n_scale, 0.0)
        return psi.astype(np.complex64), n.astype(np.float32)


def default_normaliser() -> Normaliser:
    p = PhysicsConstants()
    return Normaliser(psi_scale=math.sqrt(p.gamma_C / p.g_C) * 0.1, n_scale=p.gamma_C / p.R)


def generate(members, dt=0.05, rtol=1e-6, max_time=4000.0, log=print):
    """Cold-start steady states; returns (features, targets, converged mask, result)."""
    t0 = time.perf_counter()
    sw = PackedSweep(members, reservoir="single", precision="single", dt=dt, check_every=1.0, rtol=rtol,
                     max_time=max_time, progress=log)
    res = sw.run()
    conv = np.isfinite(res.stop_times)
    norm = default_normaliser()
    y = norm.targets(res.final_psi, [r[0] for r in res.final_res])
    log(f"generated {len(members)} members ({conv.sum()} converged) in {time.perf_counter() - t0:.1f} s")
    return features(members), y, conv, res


def train(x, y, epochs=40, batch=16, lr=2e-3, seed=0, log=print, cfg=None, time_limit=None, init=None):
    """Train on (4, N, H, W) / (3, N, H, W) arrays, from ``init`` (a parameter dict) if given;
    returns (params as numpy dict, history, cfg)."""
    import torch
    from .fno import FNOConfig, FNOModule
    cfg = cfg or FNOConfig(H=x.shape[2], W=x.shape[3])
    torch.manual_seed(seed)
    dev = "cuda"
    X = torch.tensor(x, device=dev)
    Y = torch.tensor(y, device=dev)
    n = X.shape[1]
    mod = FNOModule(cfg, seed=seed, device=dev)
    if init is not None:
        with torch.no_grad():
            for k, v in init.items():
                mod.params[k].copy_(torch.tensor(v))
    opt = torch.optim.AdamW(mod.parameters(), lr=lr, weight_decay=1e-5)
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=total, pct_start=0.1)
    hist = []
    g = torch.Generator(device="cpu").manual_seed(seed)
    t0 = time.perf_counter()
    next_log = t0 + 30.0
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g)
        acc = 0.0
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch].to(dev)
            xb, yb = _augment(X[:, idx], Y[:, idx], g)
            out = mod(xb)
            loss = ((out - yb) ** 2).mean()
            opt.zero_grad(set_to_none=True)
            loss.backward()
            torch.nn.utils.clip_grad_norm_(mod.parameters(), 1.0)
            opt.step()
            sched.step()
            acc += float(loss.detach())
            step += 1
            if time.perf_counter() > next_log:
                next_log = time.perf_counter() + 30.0
                log(f"epoch {ep + 1} step {s + 1}/{steps_per_epoch} loss {acc / (s + 1):.4e} "
                    f"({time.perf_counter() - t0:.0f} s)")
        hist.append(acc / steps_per_epoch)
        log(f"epoch {ep + 1}/{epochs} loss {hist[-1]:.4e} ({time.perf_counter() - t0:.0f} s)")
        if time_limit and time.perf_counter() - t0 > time_limit:
            log("training stopped at the time limit")
            break
    params = {k: v.detach().cpu().numpy() for k, v in mod.params.items()}
    return params, hist, cfg


def _augment(xb, yb, gen):
    """One random symmetry of the square (flips, transpose) and one random circular shift,
    applied to inputs and targets alike."""
    import torch
    k = int(torch.randint(0, 8, (1,), generator=gen))
    sy, sx = (int(v) for v in torch.randint(0, xb.shape[2], (2,), generator=gen))
    out = []
    for t in (xb, yb):
        if k & 1:
            t = torch.flip(t, dims=(3,))
        if k & 2:
            t = torch.flip(t, dims=(2,))
        if k & 4:
            t = t.transpose(2, 3)
        out.append(torch.roll(t, shifts=(sy, sx), dims=(2, 3)).contiguous())
    return out[0], out[1]


def predict(params, x, cfg=None, batch=16):
    import torch
    from .fno import FNOConfig, FNOModule
    cfg = cfg or FNOConfig(H=x.shape[2], W=x.shape[3])
    mod = FNOModule(cfg, device="cuda")
    with torch.no_grad():
        for k, v in params.items():
            mod.params[k].copy_(torch.tensor(v))
        X = torch.tensor(x, device="cuda")
        outs = []
        for s in range(0, X.shape[1], batch):
            outs.append(mod.net.forward(X[:, s:s + batch].contiguous(), {k: mod.params[k] for k in mod.names}))
        torch.cuda.synchronize()
        return torch.cat(outs, dim=1).cpu().numpy()


def predict_states(params, members, cfg=None, norm: Normaliser | None = None, batch=16):
    """The warm-start pipeline on the device: features (as ``features``) built from the
    envelopes on the GPU, the FNO forward, and the denormalised fields with each member's peak
    density (as ``Normaliser.fields``, single precision).  Returns host arrays
    (psi (B, H, W) complex64, n (B, H, W) float32, peak (B,) float32) for ``warm_start``."""
    import torch
    from .fno import FNOConfig, FNOModule
    norm = norm or default_normaliser()
    H, W = members[0].ny, members[0].nx
    cfg = cfg or FNOConfig(H=H, W=W)
    ratio = np.asarray([m.P0 / threshold(m.physics, "single") for m in members])
    gam = np.asarray([m.physics.gamma_C / 0.083 for m in members], dtype=np.float32)
    gc = np.asarray([m.physics.g_C / 0.001 for m in members], dtype=np.float32)
    env = np.stack([np.asarray(m.envelope) for m in members])
    mod = FNOModule(cfg, device="cuda")
    with torch.no_grad():
        for k, v in params.items():
            mod.params[k].copy_(torch.tensor(v))
        dev = torch.device("cuda")
        E = torch.tensor(env, device=dev)
        R = torch.tensor(ratio, device=dev)
        B = len(members)
        X = torch.empty((IN_FEATURES, B, H, W), dtype=torch.float32, device=dev)
        X[0] = (E * R[:, None, None]).to(torch.float32)
        X[1] = R.to(torch.float32)[:, None, None]
        X[2] = torch.tensor(gam, device=dev)[:, None, None]
        X[3] = torch.tensor(gc, device=dev)[:, None, None]
        outs = []
        for s0 in range(0, B, batch):
            outs.append(mod.net.forward(X[:, s0:s0 + batch].contiguous(), {k: mod.params[k] for k in mod.names}))
        Y = torch.cat(outs, dim=1)
        re = Y[0] * np.float32(norm.psi_scale)
        im = Y[1] * np.float32(norm.psi_scale)
        n = torch.clamp(Y[2] * np.float32(norm.n_scale), min=0.0)
        peak = (re * re + im * im).reshape(B, -1).amax(dim=1)
        psi = torch.complex(re, im)
        torch.cuda.synchronize()
        return psi.cpu().numpy(), n.cpu().numpy(), peak.cpu().numpy()


def warm_start(members, y_pred, norm: Normaliser | None = None, states=None):
    """Members whose initial state is the predicted (psi, n_R), from the network output
    ``y_pred`` or from the ``states`` of ``predict_states``.

    The cold start's noise field (RMS 1e-3) is not added: the prediction's own error is larger,
    so it seeds any symmetry breaking the dynamics needs.  A member predicted to hold no
    condensate (peak density below ``ZERO_STATE_PEAK``) starts from psi = 0 exactly; only the
    cold start of the same member can confirm that prediction."""
    norm = norm or default_normaliser()
    B = len(members)
    if states is not None:
        psi_all, n_all, peak = states
    else:
        # all members at once, with the arithmetic of Normaliser.fields (single precision)
        re = y_pred[0, :B] * np.float32(norm.psi_scale)
        im = y_pred[1, :B] * np.float32(norm.psi_scale)
        n_all = np.maximum(y_pred[2, :B] * np.float32(norm.n_scale), np.float32(0.0))
        peak = (re * re + im * im).reshape(B, -1).max(axis=1)
        psi_all = np.empty(re.shape, np.complex64)
        psi_all.real = re
        psi_all.imag = im
    out = []
    for b, m in enumerate(members):
        label = dict(m.label)
        if float(peak[b]) < ZERO_STATE_PEAK:
            psi = np.zeros((m.ny, m.nx), np.complex64)
            label["warm"] = "zero"
        else:
            psi = psi_all[b]
            label["warm"] = "predicted"
        out.append(SweepMember(m.nx, m.ny, m.lx, m.ly, m.physics, m.envelope, m.P0, seed=m.seed,
                               psi0=psi, res0=(n_all[b],), label=label))
    return out


def save_weights(path, params, cfg, norm: Normaliser, meta=None):
    arrays = {f"param/{k}": v.astype(np.float32) for k, v in params.items()}
    arrays["cfg"] = np.asarray([cfg.H, cfg.W, cfg.width, cfg.layers, cfg.mx, cfg.my, cfg.fin, cfg.hidden, cfg.fout])
    arrays["norm"] = np.asarray([norm.psi_scale, norm.n_scale])
    if meta:
        for k, v in meta.items():
            arrays[f"meta/{k}"] = np.asarray(v)
    np.savez_compressed(path, **arrays)


def load_weights(path):
    from .fno import FNOConfig
    with np.load(path, allow_pickle=False) as z:
        c = [int(v) for v in z["cfg"]]
        cfg = FNOConfig(H=c[0], W=c[1], width=c[2], layers=c[3], modes_x=c[4], modes_y=c[5], in_features=c[6],
                        hidden=c[7], out_features=c[8])
        params = {k[len("param/"):]: z[k] for k in z.files if k.startswith("param/")}
        norm = Normaliser(float(z["norm"][0]), float(z["norm"][1]))
    return params, cfg, norm
