"""torch.autograd.Function over the C ABI (used by the test harness).

y = A(vals) x with A an AXT NOC matrix in column-map form.  Backward:
  x_bar    = A^T y_bar   -- spmv_axt_gather_permuted on the transposed tile map,
                           reading A's current values through the slot map
                           (or spmv_permute_vals + the plain product, where
                           tune_adjoint finds that faster)
  vals_bar = y_bar[row(s)] * x[col(s)] per slot -- spmv_axt_grad_vals
Both are hand-written gfx950 kernels; torch only holds the tensors.

AxtSpMVPair holds A's values twice, in A's slot order (read by the forward)
and in A^T's (read by x_bar = A^T y_bar), as two leaves: its backward returns
the value gradient in both orders -- the A^T-order one written by the adjoint
kernel in the same pass (spmv_axt_gather_grad) -- so an elementwise optimiser
updates the two copies alike, entry for entry, and the adjoint never reads
A's values through the slot map.  (A global-norm computation over the
parameters counts the pair twice.)
"""
from __future__ import annotations

import torch

from . import device as D


class AxtOperator:
    """Structure of A (slot order, column map, plan) and of A^T."""

    def __init__(self, eng: D.Engine, m, th: int = 1, val_type: int = D.F64):
        from . import host as H
        self.eng, self.m = eng, m
        zero = __import__("numpy").zeros(m.nrows)
        self.a = H.AXT(m, zero, th=th, want_acol=True)
        self.fwd = D.DeviceAXTGather(eng, self.a, val_type, val_type)
        self.adj = D.AxtAdjoint(eng, m, self.fwd)
        self.dtype = torch.float64 if val_type == D.F64 else torch.float32

    def tune_adjoint(self) -> dict:
        """Picks, for this matrix, how backward's x_bar takes the current
        values (through the slot map, or a refresh of A^T's copy first)."""
        return self.adj.tune_values_path()

    def tune_placement(self) -> dict:
        """Picks, for this matrix, the plan placement (row groups pinned to
        XCDs or not; same bits) of the forward and of x_bar = A^T y_bar."""
        return dict(forward=self.fwd.tune_placement(k=1), adjoint=self.adj.gT.tune_placement(k=1))

    def initial_values(self):
        """The matrix values in slot order (a leaf tensor to optimise)."""
        return self.fwd.aval.clone()

    def initial_values_pair(self):
        """The matrix values in A's and in A^T's slot order (two leaves for
        AxtSpMVPair)."""
        self.adj.refresh(self.fwd.aval)
        return self.fwd.aval.clone(), self.adj.gT.aval.clone()


class AxtSpMV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, x, op: AxtOperator):
        vals, x = vals.contiguous(), x.contiguous()
        op.eng.follow_current_stream()
        y = torch.empty(op.m.nrows, dtype=op.dtype, device=x.device)
        op.fwd.matvec(x, y, vals=vals)
        ctx.save_for_backward(vals, x)
        ctx.op = op
        return y

    @staticmethod
    def backward(ctx, ybar):
        vals, x = ctx.saved_tensors
        op = ctx.op
        ybar = ybar.contiguous()
        op.eng.follow_current_stream()
        gx = gv = None
        if ctx.needs_input_grad[1]:
            gx = torch.empty_like(x)
            op.adj.adjoint(ybar, gx, vals=vals)  # follows A's current values
        if ctx.needs_input_grad[0]:
            gv = torch.empty_like(vals)
            op.fwd.grad_vals(x, ybar, gv)
        return gv, gx, None


class AxtSpMVPair(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, vals_t, x, op: AxtOperator):
        vals, vals_t, x = vals.contiguous(), vals_t.contiguous(), x.contiguous()
        op.eng.follow_current_stream()
        y = torch.empty(op.m.nrows, dtype=op.dtype, device=x.device)
        op.fwd.matvec(x, y, vals=vals)
        ctx.save_for_backward(vals, vals_t, x)
        ctx.op = op
        return y

    @staticmethod
    def backward(ctx, ybar):
        vals, vals_t, x = ctx.saved_tensors
        op = ctx.op
        ybar = ybar.contiguous()
        op.eng.follow_current_stream()
        gv = gvt = gx = None
        if ctx.needs_input_grad[0]:
            gv = torch.empty_like(vals)
            op.fwd.grad_vals(x, ybar, gv)
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            gx, gvt = torch.empty_like(x), torch.empty_like(vals_t)
            op.adj.adjoint_grad(ybar, gx, x, gvt, vals_t=vals_t)
            gx = gx if ctx.needs_input_grad[2] else None
            gvt = gvt if ctx.needs_input_grad[1] else None
        return gv, gvt, gx, None


def spmv(vals, x, op: AxtOperator):
    return AxtSpMV.apply(vals, x, op)


def spmv_pair(vals, vals_t, x, op: AxtOperator):
    return AxtSpMVPair.apply(vals, vals_t, x, op)
