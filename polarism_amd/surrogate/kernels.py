"""Thin wrappers that run the surrogate's HIP kernels on torch-owned memory.

torch only allocates the tensors and supplies the current stream; every
product and elementwise op of the network runs in libpgpe_hip.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _native as N


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _ptr(t):
    return None if t is None else t.data_ptr()


def gemm(A, B, C_out, M, N_, K, *, batch=1, sa=(None, None, 0), sb=(None, None, 0), sc=(None, None, 0),
         alpha=1.0, beta=0.0, bias=None, bias_mode=0, residual=None, srb=0, gelu=False):
    """C[b](m, n) = alpha sum_k A[b](m, k) B[b](k, n) + beta C + bias + residual.

    ``sa = (sam, sak, sab)``, ``sb = (sbk, sbn, sbb)``, ``sc = (scm, scn, scb)`` in elements;
    A and B are bf16 tensors, C is fp32 or bf16.
    """
    if A.dtype != torch.bfloat16 or B.dtype != torch.bfloat16:
        raise TypeError("gemm operands must be bfloat16")
    a = N.GemmArgs()
    a.A, a.B, a.C = A.data_ptr(), B.data_ptr(), C_out.data_ptr()
    a.bias = _ptr(bias)
    a.residual = _ptr(residual)
    a.M, a.N, a.K, a.batch = int(M), int(N_), int(K), int(batch)
    a.sam, a.sak, a.sab = sa
    a.sbk, a.sbn, a.sbb = sb
    a.scm, a.scn, a.scb = sc
    a.srb = srb
    a.alpha, a.beta = float(alpha), float(beta)
    a.bias_mode = bias_mode
    a.gelu = 1 if gelu else 0
    a.out_bf16 = 1 if C_out.dtype == torch.bfloat16 else 0
    N.check(N.lib().pgpe_gemm_bf16(C.byref(a), _stream()), "pgpe_gemm_bf16")


def to_bf16(x, out=None):
    out = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device) if out is None else out
    N.call("pgpe_f32_to_bf16", x.data_ptr(), out.data_ptr(), x.numel(), _stream())
    return out


def gelu_fwd(x, y=None, yb=None):
    N.call("pgpe_gelu_fwd", x.data_ptr(), _ptr(y), _ptr(yb), x.numel(), _stream())


def gelu_bwd(x, dy, dx=None, dxb=None):
    N.call("pgpe_gelu_bwd", x.data_ptr(), dy.data_ptr(), _ptr(dx), _ptr(dxb), x.numel(), _stream())


def colsum(x, out, rows, cols, row_stride, col_stride):
    N.call("pgpe_colsum", x.data_ptr(), out.data_ptr(), rows, cols, row_stride, col_stride, _stream())


def mode_scale(vr, vi, dr, di, ur=None, ui=None, urb=None, uib=None, conj=False):
    """u = d v per mode and channel; v, u planar (my, C, B, mx); d (C, my, mx)."""
    my, C_, B, mx = vr.shape
    N.call("pgpe_mode_scale", vr.data_ptr(), vi.data_ptr(), dr.data_ptr(), di.data_ptr(), _ptr(ur), _ptr(ui),
           _ptr(urb), _ptr(uib), my, C_, B, mx, 1 if conj else 0, _stream())


def mode_scale_grad(vr, vi, dur, dui, gdr, gdi):
    my, C_, B, mx = vr.shape
    N.call("pgpe_mode_scale_grad", vr.data_ptr(), vi.data_ptr(), dur.data_ptr(), dui.data_ptr(), gdr.data_ptr(),
           gdi.data_ptr(), my, C_, B, mx, _stream())
