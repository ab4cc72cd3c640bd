"""Emulated DGEMM as a differentiable torch operation.

``ozaki_matmul(A, B)`` computes C = A B with the Ozaki-scheme kernels of libozaie.so;
its backward runs the two gradient products as emulated GEMMs as well:

    dA = dC B^T    (M x N) (N x K): contraction over N
    dB = A^T dC    (K x M) (M x N): contraction over M

The transposed operands are not materialised and no forward slices are kept: each
backward product re-splits its operands along its own contraction axis (the split
kernel reads A^T / B^T through strides), so every product uses row/column exponents
of the operand it actually contracts.  ``crt=True`` runs all three products in the CRT
mode (``slices`` moduli, ``bits`` = 0 for the largest exact bit count).
"""
import torch

from . import Plan, ROUND_NEAREST, ROUND_TRUNCATE, _unit_last_axis

_plans = {}


def _row_major(x):
    return _unit_last_axis(x)


def _plan(M, N, K, bits, slices, rounding, crt=False):
    # A plan owns its intermediate buffers (exponents, slices or residues, partial sums), so
    # products enqueued on different streams must not share one: the key holds the stream.
    key = (M, N, K, bits, slices, rounding, crt, torch.cuda.current_device(),
           torch.cuda.current_stream().cuda_stream)
    p = _plans.get(key)
    if p is None:
        p = Plan(M, N, K, bits, slices, rounding, crt=crt)
        _plans[key] = p
    return p


def clear_plan_cache():
    """Free the device buffers of every cached plan."""
    torch.cuda.synchronize()
    for p in _plans.values():
        p.close()
    _plans.clear()


class OzakiMatmul(torch.autograd.Function):
    """C = A @ B on float64 device tensors, emulated with b-bit int8 slices."""

    @staticmethod
    def forward(ctx, A, B, bits=7, slices=9, rounding=ROUND_NEAREST, crt=False):
        if A.dim() != 2 or B.dim() != 2 or A.shape[1] != B.shape[0]:
            raise ValueError(f"shapes {tuple(A.shape)} @ {tuple(B.shape)} do not form a matrix product")
        if A.dtype != torch.float64 or B.dtype != torch.float64 or not (A.is_cuda and B.is_cuda):
            raise ValueError("ozaki_matmul takes float64 device tensors")
        A = A if _row_major(A) else A.contiguous()
        B = B if _row_major(B) else B.contiguous()
        M, K = A.shape
        N = B.shape[1]
        ctx.save_for_backward(A, B)
        ctx.cfg = (bits, slices, rounding, crt)
        return _plan(M, N, K, bits, slices, rounding, crt)(A, B)

    @staticmethod
    def backward(ctx, dC):
        A, B = ctx.saved_tensors
        bits, slices, rounding, crt = ctx.cfg
        M, K = A.shape
        N = B.shape[1]
        dC = dC if (dC.dtype == torch.float64 and _row_major(dC)) else dC.to(torch.float64).contiguous()
        dA = dB = None
        if ctx.needs_input_grad[0]:
            # dA (M x K) = dC (M x N) . B^T; B is stored K x N = (N x K)^T -> transB
            dA = _plan(M, K, N, bits, slices, rounding, crt)(dC, B, transB=True)
        if ctx.needs_input_grad[1]:
            # dB (K x N) = A^T (K x M) . dC (M x N); A is stored M x K = (K x M)^T -> transA
            dB = _plan(K, N, M, bits, slices, rounding, crt)(A, dC, transA=True)
        return dA, dB, None, None, None, None


def ozaki_matmul(A, B, bits=7, slices=9, rounding=ROUND_NEAREST, crt=False):
    """Differentiable emulated C = A @ B (float64 device tensors); crt=True: CRT mode with
    `slices` moduli (use bits=0, slices=16)."""
    return OzakiMatmul.apply(A, B, bits, slices, rounding, crt)


__all__ = ["OzakiMatmul", "ozaki_matmul", "clear_plan_cache", "ROUND_TRUNCATE", "ROUND_NEAREST"]
