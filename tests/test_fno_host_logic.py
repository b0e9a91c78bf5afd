"""FNO host logic (strides, layouts, backward algebra) on the CPU.

The HIP entry points are replaced by torch emulations of their documented
semantics (pgpe_gemm_bf16's strided batched product with bf16 operands, the
GELU and column-sum kernels), so the composition in polarism_amd/surrogate/fno.py
is checked against the float64 autograd reference without a GPU.
"""
import math

import numpy as np
import pytest
import torch

from polarism_amd.surrogate import fno as F


def _gemm(A, B, C, M, N_, K, *, batch=1, sa=(None, None, 0), sb=(None, None, 0), sc=(None, None, 0), alpha=1.0,
          beta=0.0, bias=None, bias_mode=0, residual=None, srb=0, gelu=False):
    assert A.dtype == torch.bfloat16 and B.dtype == torch.bfloat16
    Av = torch.as_strided(A, (batch, M, K), (sa[2], sa[0], sa[1])).double()
    Bv = torch.as_strided(B, (batch, K, N_), (sb[2], sb[0], sb[1])).double()
    Cv = torch.as_strided(C, (batch, M, N_), (sc[2], sc[0], sc[1]))
    v = alpha * torch.bmm(Av, Bv)
    if beta:
        v = v + beta * Cv.double()
    if bias_mode == 1:
        v = v + bias.double()[None, :, None]
    elif bias_mode == 2:
        v = v + bias.double()[None, None, :]
    if residual is not None:
        v = v + torch.as_strided(residual, (batch, M, N_), (srb, sc[0], sc[1])).double()
    if gelu:
        v = torch.nn.functional.gelu(v)
    Cv.copy_(v.to(C.dtype))


def _to_bf16(x, out=None):
    y = x.to(torch.bfloat16)
    if out is not None:
        out.copy_(y)
        return out
    return y


def _gelu_fwd(x, y=None, yb=None):
    v = torch.nn.functional.gelu(x)
    if y is not None:
        y.copy_(v)
    if yb is not None:
        yb.copy_(v.to(torch.bfloat16))


def _gelu_bwd(x, dy, dx=None, dxb=None):
    xx = x.double()
    cdf = 0.5 * (1 + torch.erf(xx / math.sqrt(2)))
    pdf = torch.exp(-0.5 * xx * xx) / math.sqrt(2 * math.pi)
    g = (dy.double() * (cdf + xx * pdf)).to(torch.float32)
    if dx is not None:
        dx.copy_(g)
    if dxb is not None:
        dxb.copy_(g.to(torch.bfloat16))


def _colsum(x, out, rows, cols, row_stride, col_stride):
    v = torch.as_strided(x, (rows, cols), (row_stride, col_stride)).double().sum(dim=0)
    out.copy_(v.float())


def _mode_scale(vr, vi, dr, di, ur=None, ui=None, urb=None, uib=None, conj=False):
    d_r = dr.permute(1, 0, 2).unsqueeze(2).double()
    d_i = di.permute(1, 0, 2).unsqueeze(2).double() * (-1.0 if conj else 1.0)
    re = d_r * vr.double() - d_i * vi.double()
    im = d_r * vi.double() + d_i * vr.double()
    for dst, val in ((ur, re), (ui, im), (urb, re), (uib, im)):
        if dst is not None:
            dst.copy_(val.to(dst.dtype))


def _mode_scale_grad(vr, vi, dur, dui, gdr, gdi):
    a, b, c, d = (t.double() for t in (vr, vi, dur, dui))
    gdr.copy_((a * c + b * d).sum(dim=2).permute(1, 0, 2).float())
    gdi.copy_((a * d - b * c).sum(dim=2).permute(1, 0, 2).float())


@pytest.fixture()
def cpu_fno(monkeypatch):
    monkeypatch.setattr(F, "gemm", _gemm)
    monkeypatch.setattr(F, "to_bf16", _to_bf16)
    monkeypatch.setattr(F, "gelu_fwd", _gelu_fwd)
    monkeypatch.setattr(F, "gelu_bwd", _gelu_bwd)
    monkeypatch.setattr(F, "colsum", _colsum)
    monkeypatch.setattr(F, "mode_scale", _mode_scale)
    monkeypatch.setattr(F, "mode_scale_grad", _mode_scale_grad)
    cfg = F.FNOConfig(H=16, W=24, width=8, layers=4, modes_x=6, modes_y=8, hidden=12)
    mod = F.FNOModule(cfg, seed=3, device="cpu")
    return cfg, mod


def test_forward_and_gradients_match_float64_reference(cpu_fno):
    from tests.test_gpu_fno import ref_forward
    cfg, mod = cpu_fno
    torch.manual_seed(0)
    x = torch.randn(cfg.fin, 2, cfg.H, cfg.W)
    target = torch.randn(cfg.fout, 2, cfg.H, cfg.W)
    out = mod(x)
    loss = ((out - target) ** 2).mean()
    loss.backward()
    p64 = {k: v.detach().double().requires_grad_(True) for k, v in mod.params.items()}
    ref = ref_forward(x.double(), p64, cfg)
    assert ((out.double() - ref).norm() / ref.norm()).item() < 3e-2
    ((ref - target.double()) ** 2).mean().backward()
    for k, prm in mod.params.items():
        g, r = prm.grad.double(), p64[k].grad
        cos = torch.nn.functional.cosine_similarity(g.flatten(), r.flatten(), dim=0).item()
        rel = ((g - r).norm() / r.norm()).item()
        assert cos > 0.995 and rel < 0.1, (k, cos, rel)
