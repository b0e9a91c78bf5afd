"""Fourier neural operator for steady-state warm starts, on hand-written kernels.

Maps a pump map plus three scalars (P0 / P_th, gamma_C, g_C), given as four
input channels on an H x W grid, to the steady state (Re psi, Im psi, n_R).  Architecture (width 64, 4 Fourier
layers, 32 x 32 kept modes: kx in [0, 32), ky in [-16, 16)):

    v0 = Lift(features)                                   pointwise 4 -> 64
    v_{l+1} = GELU(Wp_l v_l + b_l + S_l v_l)  (no GELU after the last layer)
    S_l v   = real inverse DFT of  M_l[ky] (d_l[ky, kx] * DFT(v)[ky, kx])
    out     = Q2 GELU(Q1 v_4 + q1) + q2                    pointwise 64 -> 128 -> 3

The spectral weights are factorised per kept mode: a complex per-mode,
per-channel scale d_l (64 x 32 x 32) followed by a complex 64 x 64 channel
mixing M_l shared by the modes of one ky row.  That is 1.60 M real
parameters in total (dense per-mode 64 x 64 complex weights would need 67 M).

Every product is a bf16 MFMA GEMM with fp32 accumulation
(``pgpe_gemm_bf16``): the truncated DFTs are GEMMs against cos / sin
matrices; complex products are four real GEMMs.  Activations are laid out
channel-major ``[C, B, H, W]`` so the pointwise layers are single GEMMs over
all pixels of the batch.  ``FNOFunction`` (a ``torch.autograd.Function``)
runs the forward and the hand-derived backward on these kernels; PyTorch only
owns the memory and runs the optimizer.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .kernels import colsum, gelu_bwd, gelu_fwd, gemm, mode_scale, mode_scale_grad, to_bf16

BF = torch.bfloat16


class FNOConfig:
    def __init__(self, H=256, W=256, width=64, layers=4, modes_x=32, modes_y=32, in_features=4, hidden=128,
                 out_features=3):
        self.H, self.W, self.width, self.layers = H, W, width, layers
        self.mx, self.my = modes_x, modes_y
        self.fin, self.hidden, self.fout = in_features, hidden, out_features


def init_params(cfg: FNOConfig, seed=0, device="cuda"):
    """fp32 master parameters (dict of tensors)."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    C = cfg.width

    def u(shape, fan_in):
        bound = 1.0 / math.sqrt(fan_in)
        return ((torch.rand(shape, generator=g) * 2 - 1) * bound).to(device)

    p = {"lift_w": u((C, cfg.fin), cfg.fin), "lift_b": u((C,), cfg.fin)}
    for l in range(cfg.layers):
        p[f"d{l}_re"] = torch.ones((C, cfg.my, cfg.mx)).to(device) + u((C, cfg.my, cfg.mx), 16.0)
        p[f"d{l}_im"] = u((C, cfg.my, cfg.mx), 16.0)
        p[f"m{l}_re"] = u((cfg.my, C, C), C * 2)
        p[f"m{l}_im"] = u((cfg.my, C, C), C * 2)
        p[f"w{l}"] = u((C, C), C)
        p[f"b{l}"] = u((C,), C)
    p["q1_w"] = u((cfg.hidden, C), C)
    p["q1_b"] = u((cfg.hidden,), C)
    p["q2_w"] = u((cfg.fout, cfg.hidden), cfg.hidden)
    p["q2_b"] = u((cfg.fout,), cfg.hidden)
    return p


def param_count(cfg: FNOConfig) -> int:
    C = cfg.width
    per_layer = 2 * C * cfg.my * cfg.mx + 2 * cfg.my * C * C + C * C + C
    return (C * cfg.fin + C) + cfg.layers * per_layer + (cfg.hidden * C + cfg.hidden) + (cfg.fout * cfg.hidden + cfg.fout)


class DFTBasis:
    """bf16 cos / sin matrices of the truncated transforms for an H x W grid."""

    def __init__(self, cfg: FNOConfig, device="cuda"):
        H, W, mx, my = cfg.H, cfg.W, cfg.mx, cfg.my
        x = np.arange(W)
        y = np.arange(H)
        kx = np.arange(mx)
        ky = np.concatenate([np.arange(my // 2), np.arange(-(my // 2), 0)])
        ax = 2 * np.pi * np.outer(x, kx) / W            # (W, mx)
        ay = 2 * np.pi * np.outer(ky, y) / H            # (my, H)
        t = lambda a: torch.tensor(a, dtype=torch.float32).to(device).to(BF).contiguous()
        # forward: V1 = v Cx - i v Sx ; Vhat = (Cy - i Sy) V1
        self.cx, self.sx = t(np.cos(ax)), t(np.sin(ax))                    # (W, mx)
        self.cy, self.sy = t(np.cos(ay)), t(np.sin(ay))                    # (my, H)
        # inverse along y: Z = (Ey_c + i Ey_s) Wk with E = exp(+i ky y) : (H, my)
        self.ecy, self.esy = t(np.cos(ay.T)), t(np.sin(ay.T))
        # inverse along x (real output, half spectrum): s = Zr Bc - Zi Bs, with weights 1 (kx=0), 2 (kx>0), / (H W)
        w = np.where(kx == 0, 1.0, 2.0)[:, None] / (H * W)
        self.bxc = t(w * np.cos(ax.T))                                      # (mx, W)
        self.bxs = t(w * np.sin(ax.T))
        self.ky = ky
        self.kx = kx


class FNO:
    """Forward / backward of the network on the HIP kernels (batch B, channel-major)."""

    def __init__(self, cfg: FNOConfig, device="cuda"):
        self.cfg = cfg
        self.basis = DFTBasis(cfg, device)
        self.device = device

    # ---------------------------------------------------------------- helpers
    def _pw(self, Wb, xb, bias, out, rows, n, k, residual=None, gelu=False):
        """out (rows x n) = W (rows x k) x (k x n) + bias[rows] (+ residual)."""
        gemm(Wb, xb, out, rows, n, k, sa=(k, 1, 0), sb=(n, 1, 0), sc=(n, 1, 0), bias=bias,
             bias_mode=1 if bias is not None else 0, residual=residual, gelu=gelu)

    def _spectral_fwd(self, l, vb, p, B, cache):
        """Spectral branch S_l v for v (C, B, H, W) given as bf16 vb; returns fp32 (C, B, H, W)."""
        cfg, bs = self.cfg, self.basis
        C, H, W, mx, my = cfg.width, cfg.H, cfg.W, cfg.mx, cfg.my
        dev = vb.device
        R = C * B * H
        # along x: V1 = v Cx - i v Sx   (R x mx)
        v1r = torch.empty(R, mx, device=dev)
        v1i = torch.empty(R, mx, device=dev)
        gemm(vb, bs.cx, v1r, R, mx, W, sa=(W, 1, 0), sb=(mx, 1, 0), sc=(mx, 1, 0))
        gemm(vb, bs.sx, v1i, R, mx, W, sa=(W, 1, 0), sb=(mx, 1, 0), sc=(mx, 1, 0), alpha=-1.0)
        v1rb, v1ib = to_bf16(v1r), to_bf16(v1i)
        # along y, per (c, b): Vhat = (Cy - i Sy) V1, written as [ky][c][b][kx]
        CB = C * B
        vr = torch.empty(my, C, B, mx, device=dev)
        vi = torch.empty(my, C, B, mx, device=dev)
        sa_y = (H, 1, 0)
        sb_y = (mx, 1, H * mx)
        sc_y = (CB * mx, 1, mx)
        gemm(bs.cy, v1rb, vr, my, mx, H, batch=CB, sa=sa_y, sb=sb_y, sc=sc_y)
        gemm(bs.sy, v1ib, vr, my, mx, H, batch=CB, sa=sa_y, sb=sb_y, sc=sc_y, beta=1.0)
        gemm(bs.cy, v1ib, vi, my, mx, H, batch=CB, sa=sa_y, sb=sb_y, sc=sc_y)
        gemm(bs.sy, v1rb, vi, my, mx, H, batch=CB, sa=sa_y, sb=sb_y, sc=sc_y, alpha=-1.0, beta=1.0)
        # per-mode channel scale: U = d * Vhat  (d laid out (C, my, mx) -> broadcast over b)
        urb = torch.empty(my, C, B, mx, device=dev, dtype=BF)
        uib = torch.empty(my, C, B, mx, device=dev, dtype=BF)
        mode_scale(vr, vi, p[f"d{l}_re"].contiguous(), p[f"d{l}_im"].contiguous(), urb=urb, uib=uib)
        # channel mixing per ky: Wk = M[ky] U[ky]   (C x (B mx))
        mr, mi = to_bf16(p[f"m{l}_re"].contiguous()), to_bf16(p[f"m{l}_im"].contiguous())
        Nn = B * mx
        wr = torch.empty(my, C, B, mx, device=dev)
        wi = torch.empty(my, C, B, mx, device=dev)
        sa_m = (C, 1, C * C)
        sb_m = (Nn, 1, C * Nn)
        sc_m = (Nn, 1, C * Nn)
        gemm(mr, urb, wr, C, Nn, C, batch=my, sa=sa_m, sb=sb_m, sc=sc_m)
        gemm(mi, uib, wr, C, Nn, C, batch=my, sa=sa_m, sb=sb_m, sc=sc_m, alpha=-1.0, beta=1.0)
        gemm(mr, uib, wi, C, Nn, C, batch=my, sa=sa_m, sb=sb_m, sc=sc_m)
        gemm(mi, urb, wi, C, Nn, C, batch=my, sa=sa_m, sb=sb_m, sc=sc_m, beta=1.0)
        wrb, wib = to_bf16(wr), to_bf16(wi)
        # inverse along y per (o, b): Z = (Ec + i Es) Wk  -> (C, B, H, mx)
        zr = torch.empty(C, B, H, mx, device=dev)
        zi = torch.empty(C, B, H, mx, device=dev)
        sa_e = (my, 1, 0)
        sb_e = (CB * mx, 1, mx)
        sc_e = (mx, 1, H * mx)
        gemm(bs.ecy, wrb, zr, H, mx, my, batch=CB, sa=sa_e, sb=sb_e, sc=sc_e)
        gemm(bs.esy, wib, zr, H, mx, my, batch=CB, sa=sa_e, sb=sb_e, sc=sc_e, alpha=-1.0, beta=1.0)
        gemm(bs.ecy, wib, zi, H, mx, my, batch=CB, sa=sa_e, sb=sb_e, sc=sc_e)
        gemm(bs.esy, wrb, zi, H, mx, my, batch=CB, sa=sa_e, sb=sb_e, sc=sc_e, beta=1.0)
        zrb, zib = to_bf16(zr), to_bf16(zi)
        # inverse along x (real): s = Zr Bc - Zi Bs    (R x W)
        s = torch.empty(C, B, H, W, device=dev)
        gemm(zrb, bs.bxc, s, R, W, mx, sa=(mx, 1, 0), sb=(W, 1, 0), sc=(W, 1, 0))
        gemm(zib, bs.bxs, s, R, W, mx, sa=(mx, 1, 0), sb=(W, 1, 0), sc=(W, 1, 0), alpha=-1.0, beta=1.0)
        if cache is not None:
            cache.update({f"vr{l}": vr, f"vi{l}": vi, f"urb{l}": urb, f"uib{l}": uib, f"mr{l}": mr, f"mi{l}": mi})
        return s

    def _spectral_bwd(self, l, ds, p, B, cache, grads):
        """Backward of the spectral branch: returns dv (C, B, H, W) fp32, accumulates d and M grads."""
        cfg, bs = self.cfg, self.basis
        C, H, W, mx, my = cfg.width, cfg.H, cfg.W, cfg.mx, cfg.my
        dev = ds.device
        R, CB, Nn = C * B * H, C * B, B * mx
        dsb = to_bf16(ds)
        # s = Zr Bc - Zi Bs  ->  dZr = ds Bc^T, dZi = -ds Bs^T   (R x mx)
        dzr = torch.empty(C, B, H, mx, device=dev)
        dzi = torch.empty(C, B, H, mx, device=dev)
        gemm(dsb, bs.bxc, dzr, R, mx, W, sa=(W, 1, 0), sb=(1, W, 0), sc=(mx, 1, 0))
        gemm(dsb, bs.bxs, dzi, R, mx, W, sa=(W, 1, 0), sb=(1, W, 0), sc=(mx, 1, 0), alpha=-1.0)
        dzrb, dzib = to_bf16(dzr), to_bf16(dzi)
        # Z = (Ec + i Es) Wk -> dWr = Ec^T dZr + Es^T dZi ; dWi = Ec^T dZi - Es^T dZr   ([ky][c][b][kx])
        dwr = torch.empty(my, C, B, mx, device=dev)
        dwi = torch.empty(my, C, B, mx, device=dev)
        sa_t = (1, my, 0)                 # E^T: (my x H) from E stored (H x my)
        sb_t = (mx, 1, H * mx)
        sc_t = (CB * mx, 1, mx)
        gemm(bs.ecy, dzrb, dwr, my, mx, H, batch=CB, sa=sa_t, sb=sb_t, sc=sc_t)
        gemm(bs.esy, dzib, dwr, my, mx, H, batch=CB, sa=sa_t, sb=sb_t, sc=sc_t, beta=1.0)
        gemm(bs.ecy, dzib, dwi, my, mx, H, batch=CB, sa=sa_t, sb=sb_t, sc=sc_t)
        gemm(bs.esy, dzrb, dwi, my, mx, H, batch=CB, sa=sa_t, sb=sb_t, sc=sc_t, alpha=-1.0, beta=1.0)
        dwrb, dwib = to_bf16(dwr), to_bf16(dwi)
        # Wk = M U: dM = dW U^H (per ky, complex), dU = M^H dW
        urb, uib, mr, mi = cache[f"urb{l}"], cache[f"uib{l}"], cache[f"mr{l}"], cache[f"mi{l}"]
        dmr = torch.empty(my, C, C, device=dev)
        dmi = torch.empty(my, C, C, device=dev)
        sa_w = (Nn, 1, C * Nn)
        sb_u = (1, Nn, C * Nn)            # U^T
        sc_mm = (C, 1, C * C)
        gemm(dwrb, urb, dmr, C, C, Nn, batch=my, sa=sa_w, sb=sb_u, sc=sc_mm)
        gemm(dwib, uib, dmr, C, C, Nn, batch=my, sa=sa_w, sb=sb_u, sc=sc_mm, beta=1.0)
        gemm(dwib, urb, dmi, C, C, Nn, batch=my, sa=sa_w, sb=sb_u, sc=sc_mm)
        gemm(dwrb, uib, dmi, C, C, Nn, batch=my, sa=sa_w, sb=sb_u, sc=sc_mm, alpha=-1.0, beta=1.0)
        grads[f"m{l}_re"] = dmr
        grads[f"m{l}_im"] = dmi
        dur = torch.empty(my, C, B, mx, device=dev)
        dui = torch.empty(my, C, B, mx, device=dev)
        sa_h = (1, C, C * C)              # M^T
        sb_w = (Nn, 1, C * Nn)
        sc_u = (Nn, 1, C * Nn)
        gemm(mr, dwrb, dur, C, Nn, C, batch=my, sa=sa_h, sb=sb_w, sc=sc_u)
        gemm(mi, dwib, dur, C, Nn, C, batch=my, sa=sa_h, sb=sb_w, sc=sc_u, beta=1.0)
        gemm(mr, dwib, dui, C, Nn, C, batch=my, sa=sa_h, sb=sb_w, sc=sc_u)
        gemm(mi, dwrb, dui, C, Nn, C, batch=my, sa=sa_h, sb=sb_w, sc=sc_u, alpha=-1.0, beta=1.0)
        # U = d Vhat: dd = sum_b conj(Vhat) dU ; dVhat = conj(d) dU
        vr, vi = cache[f"vr{l}"], cache[f"vi{l}"]
        gdr = torch.empty(C, my, mx, device=dev)
        gdi = torch.empty(C, my, mx, device=dev)
        mode_scale_grad(vr, vi, dur, dui, gdr, gdi)
        grads[f"d{l}_re"] = gdr
        grads[f"d{l}_im"] = gdi
        dvrb = torch.empty(my, C, B, mx, device=dev, dtype=BF)
        dvib = torch.empty(my, C, B, mx, device=dev, dtype=BF)
        mode_scale(dur, dui, p[f"d{l}_re"].contiguous(), p[f"d{l}_im"].contiguous(), urb=dvrb, uib=dvib, conj=True)
        # Vhat = (Cy - i Sy) V1 -> dV1r = Cy^T dVr - Sy^T dVi ; dV1i = Cy^T dVi + Sy^T dVr   ((c, b) x H x mx)
        dv1r = torch.empty(C, B, H, mx, device=dev)
        dv1i = torch.empty(C, B, H, mx, device=dev)
        sa_c = (1, H, 0)                  # Cy^T: (H x my) from (my x H)
        sb_v = (CB * mx, 1, mx)
        sc_v = (mx, 1, H * mx)
        gemm(bs.cy, dvrb, dv1r, H, mx, my, batch=CB, sa=sa_c, sb=sb_v, sc=sc_v)
        gemm(bs.sy, dvib, dv1r, H, mx, my, batch=CB, sa=sa_c, sb=sb_v, sc=sc_v, alpha=-1.0, beta=1.0)
        gemm(bs.cy, dvib, dv1i, H, mx, my, batch=CB, sa=sa_c, sb=sb_v, sc=sc_v)
        gemm(bs.sy, dvrb, dv1i, H, mx, my, batch=CB, sa=sa_c, sb=sb_v, sc=sc_v, beta=1.0)
        dv1rb, dv1ib = to_bf16(dv1r), to_bf16(dv1i)
        # V1 = v Cx - i v Sx -> dv = dV1r Cx^T - dV1i Sx^T
        dv = torch.empty(C, B, H, W, device=dev)
        gemm(dv1rb, bs.cx, dv, R, W, mx, sa=(mx, 1, 0), sb=(1, mx, 0), sc=(W, 1, 0))
        gemm(dv1ib, bs.sx, dv, R, W, mx, sa=(mx, 1, 0), sb=(1, mx, 0), sc=(W, 1, 0), alpha=-1.0, beta=1.0)
        return dv

    # ---------------------------------------------------------------- network
    def forward(self, x, p, cache=None):
        """x: (fin, B, H, W) fp32 features -> (fout, B, H, W) fp32."""
        cfg = self.cfg
        C, H, W, L = cfg.width, cfg.H, cfg.W, cfg.layers
        B = x.shape[1]
        P = B * H * W
        dev = x.device
        xb = to_bf16(x.contiguous())
        v = torch.empty(C, B, H, W, device=dev)
        self._pw(to_bf16(p["lift_w"].contiguous()), xb, p["lift_b"], v, C, P, cfg.fin)
        vb = to_bf16(v)
        if cache is not None:
            cache["xb"] = xb
        for l in range(L):
            s = self._spectral_fwd(l, vb, p, B, cache)
            h = torch.empty(C, B, H, W, device=dev)
            wb = to_bf16(p[f"w{l}"].contiguous())
            self._pw(wb, vb, p[f"b{l}"], h, C, P, C, residual=s)
            if cache is not None:
                cache[f"vb{l}"] = vb
                cache[f"h{l}"] = h
                cache[f"wb{l}"] = wb
            if l < L - 1:
                vb = torch.empty(C, B, H, W, device=dev, dtype=BF)
                gelu_fwd(h, None, vb)
            else:
                vb = to_bf16(h)
        q = torch.empty(cfg.hidden, B, H, W, device=dev)
        q1b = to_bf16(p["q1_w"].contiguous())
        self._pw(q1b, vb, p["q1_b"], q, cfg.hidden, P, C)
        qb = torch.empty(cfg.hidden, B, H, W, device=dev, dtype=BF)
        gelu_fwd(q, None, qb)
        out = torch.empty(cfg.fout, B, H, W, device=dev)
        q2b = to_bf16(p["q2_w"].contiguous())
        self._pw(q2b, qb, p["q2_b"], out, cfg.fout, P, cfg.hidden)
        if cache is not None:
            cache.update(vL=vb, q=q, qb=qb, q1b=q1b, q2b=q2b)
        return out

    def backward(self, dout, p, cache):
        """Gradients of every parameter for the cached forward; dout (fout, B, H, W)."""
        cfg = self.cfg
        C, H, W, L = cfg.width, cfg.H, cfg.W, cfg.layers
        B = dout.shape[1]
        P = B * H * W
        dev = dout.device
        grads = {}
        ones = torch.ones(P, device=dev)
        doutb = to_bf16(dout.contiguous())

        def wgrad(dyb, xb, rows, k, name):
            # dW (rows x k) = dy (rows x P) x^T (P x k)
            g = torch.empty(rows, k, device=dev)
            gemm(dyb, xb, g, rows, k, P, sa=(P, 1, 0), sb=(1, P, 0), sc=(k, 1, 0))
            grads[name] = g

        def bgrad(dy, rows, name):
            g = torch.empty(rows, device=dev)
            colsum(dy, g, P, rows, 1, P)
            grads[name] = g

        # out = Q2 qb + q2
        wgrad(doutb, cache["qb"], cfg.fout, cfg.hidden, "q2_w")
        bgrad(dout, cfg.fout, "q2_b")
        dqa = torch.empty(cfg.hidden, B, H, W, device=dev)
        gemm(cache["q2b"], doutb, dqa, cfg.hidden, P, cfg.fout, sa=(1, cfg.hidden, 0), sb=(P, 1, 0), sc=(P, 1, 0))
        dq = torch.empty_like(dqa)
        dqb = torch.empty(cfg.hidden, B, H, W, device=dev, dtype=BF)
        gelu_bwd(cache["q"], dqa, dq, dqb)
        wgrad(dqb, cache["vL"], cfg.hidden, C, "q1_w")
        bgrad(dq, cfg.hidden, "q1_b")
        dv = torch.empty(C, B, H, W, device=dev)
        gemm(cache["q1b"], dqb, dv, C, P, cfg.hidden, sa=(1, C, 0), sb=(P, 1, 0), sc=(P, 1, 0))
        for l in reversed(range(L)):
            h = cache[f"h{l}"]
            if l < L - 1:
                dh = torch.empty_like(h)
                dhb = torch.empty(C, B, H, W, device=dev, dtype=BF)
                gelu_bwd(h, dv, dh, dhb)
            else:
                dh = dv
                dhb = to_bf16(dv)
            vb = cache[f"vb{l}"]
            wgrad(dhb, vb, C, C, f"w{l}")
            bgrad(dh, C, f"b{l}")
            dv_new = self._spectral_bwd(l, dh, p, B, cache, grads)
            gemm(cache[f"wb{l}"], dhb, dv_new, C, P, C, sa=(1, C, 0), sb=(P, 1, 0), sc=(P, 1, 0), beta=1.0)
            dv = dv_new
        dvb = to_bf16(dv)
        wgrad(dvb, cache["xb"], C, cfg.fin, "lift_w")
        bgrad(dv, C, "lift_b")
        del ones
        return grads


PARAM_ORDER = None


class FNOFunction(torch.autograd.Function):
    """Autograd bridge: forward and backward run on the HIP kernels."""

    @staticmethod
    def forward(ctx, net, names, x, *params):
        p = dict(zip(names, params))
        cache = {}
        out = net.forward(x, p, cache)
        ctx.net, ctx.names, ctx.cache, ctx.p = net, names, cache, p
        return out

    @staticmethod
    def backward(ctx, dout):
        grads = ctx.net.backward(dout.contiguous(), ctx.p, ctx.cache)
        ctx.cache = None
        return (None, None, None) + tuple(grads[n] for n in ctx.names)


class FNOModule(torch.nn.Module):
    def __init__(self, cfg: FNOConfig, seed=0, device="cuda"):
        super().__init__()
        self.cfg = cfg
        self.net = FNO(cfg, device)
        params = init_params(cfg, seed, device)
        self.names = tuple(params)
        self.params = torch.nn.ParameterDict({k: torch.nn.Parameter(v) for k, v in params.items()})

    def forward(self, x):
        return FNOFunction.apply(self.net, self.names, x, *[self.params[n] for n in self.names])
