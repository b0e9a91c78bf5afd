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
    A and B are bf16 tensors, C is fp32 or bf16.  A long reduction (K >= 2^16) into a
    small contiguous fp32 output (the weight gradients, K = pixels of the batch) is split
    over K into partial products reduced in a fixed order, so it spreads over the chip.
    """
    if A.dtype != torch.bfloat16 or B.dtype != torch.bfloat16:
        raise TypeError("gemm operands must be bfloat16")
    splits = _k_splits(M, N_, K, batch, sc, beta, bias, residual, gelu, C_out)
    if splits > 1:
        kc = K // splits
        part = torch.empty((splits, M, N_), dtype=torch.float32, device=C_out.device)
        _launch(A, B, part, M, N_, kc, splits, (sa[0], sa[1], kc * sa[1]), (sb[0], sb[1], kc * sb[0]),
                (N_, 1, M * N_), alpha, 0.0, None, 0, None, 0, False)
        colsum(part, C_out, splits, M * N_, M * N_, 1)
        return
    _launch(A, B, C_out, M, N_, K, batch, sa, sb, sc, alpha, beta, bias, bias_mode, residual, srb, gelu)


def _k_splits(M, N_, K, batch, sc, beta, bias, residual, gelu, C_out):
    tiles = ((M + 127) // 128) * ((N_ + 127) // 128) * batch
    if (batch != 1 or K < (1 << 16) or tiles >= 64 or beta != 0.0 or bias is not None or residual is not None
            or gelu or C_out.dtype != torch.float32 or sc[0] != N_ or sc[1] != 1):
        return 1
    splits = 1
    while splits < 256 and K % (2 * splits) == 0 and (K // (2 * splits)) % 32 == 0 and K // (2 * splits) >= 2048:
        splits *= 2
    return splits


def _launch(A, B, C_out, M, N_, K, batch, sa, sb, sc, alpha, beta, bias, bias_mode, residual, srb, gelu):
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
    """out[c] = sum_r x[r row_stride + c col_stride].  Long contiguous columns (one workgroup
    each would leave the chip idle) are first summed in ``chunks`` contiguous pieces, each its own
    workgroup, then the pieces of each column are summed: a fixed order either way."""
    chunks = 1
    if row_stride == 1 and col_stride == rows:
        while chunks < 256 and rows % (2 * chunks) == 0 and rows // (2 * chunks) >= 4096 and cols * chunks < 2048:
            chunks *= 2
    if chunks == 1:
        N.call("pgpe_colsum", x.data_ptr(), out.data_ptr(), rows, cols, row_stride, col_stride, _stream())
        return
    piece = rows // chunks
    part = torch.empty(cols * chunks, dtype=torch.float32, device=out.device)
    N.call("pgpe_colsum", x.data_ptr(), part.data_ptr(), piece, cols * chunks, 1, piece, _stream())
    N.call("pgpe_colsum", part.data_ptr(), out.data_ptr(), chunks, cols, 1, chunks, _stream())


def mode_scale(vr, vi, dr, di, ur=None, ui=None, urb=None, uib=None, conj=False):
    """u = d v per mode and channel; v, u planar (my, C, B, mx); d (C, my, mx)."""
    my, C_, B, mx = vr.shape
    N.call("pgpe_mode_scale", vr.data_ptr(), vi.data_ptr(), dr.data_ptr(), di.data_ptr(), _ptr(ur), _ptr(ui),
           _ptr(urb), _ptr(uib), my, C_, B, mx, 1 if conj else 0, _stream())


def mode_scale_grad(vr, vi, dur, dui, gdr, gdi):
    my, C_, B, mx = vr.shape
    N.call("pgpe_mode_scale_grad", vr.data_ptr(), vi.data_ptr(), dur.data_ptr(), dui.data_ptr(), gdr.data_ptr(),
           gdi.data_ptr(), my, C_, B, mx, _stream())
