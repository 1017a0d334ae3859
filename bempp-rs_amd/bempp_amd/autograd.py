"""torch.autograd wrapper of the bilinear form L(X, u, v) = v^T A(X) u.

X are the grid's vertex coordinates. The forward pass applies the operator
matrix-free on the GPU (no assembled matrix). The backward pass calls the
device adjoint kernel for dL/dX (shape gradient), and applies the operator again
for dL/du = A^T v and dL/dv = A u. Gradients follow torch's convention for
complex outputs: a real input x receives Re(conj(g) dL/dx), and a complex input z
receives g conj(dL/dz).

Implemented for the single-layer operator on one P0 space.
"""
import numpy as np
import torch


def _np(t):
    return t.detach().cpu().numpy()


class BilinearForm(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, u, v, assembler, space):
        grid = space.grid
        grid.set_points(_np(points).astype(np.float64))
        un, vn = _np(u), _np(v)
        y = assembler.apply(space, space, un)
        L = np.sum(vn * y)
        ctx.save_for_backward(points, u, v)
        ctx.assembler, ctx.space, ctx.y = assembler, space, y
        return torch.tensor(L, dtype=torch.complex128 if np.iscomplexobj(L) else torch.float64)

    @staticmethod
    def backward(ctx, g):
        points, u, v = ctx.saved_tensors
        asm, space = ctx.assembler, ctx.space
        gn = complex(g.detach().cpu().item())
        space.grid.set_points(_np(points).astype(np.float64))
        grad_points = grad_u = grad_v = None
        if ctx.needs_input_grad[0]:
            G = asm.shape_gradient(space, space, _np(u), _np(v))
            grad_points = torch.from_numpy(np.real(np.conj(gn) * G)).to(points.device, points.dtype)
        if ctx.needs_input_grad[1]:
            atv = asm.apply(space, space, _np(v), transpose=True)
            grad_u = torch.from_numpy(gn * np.conj(atv)).to(u.device, u.dtype)
        if ctx.needs_input_grad[2]:
            grad_v = torch.from_numpy(gn * np.conj(ctx.y)).to(v.device, v.dtype)
        return grad_points, grad_u, grad_v, None, None


def bilinear_form(points, u, v, assembler, space):
    """L = v^T A(points) u, differentiable in points, u and v (torch tensors)."""
    if u.shape[0] != space.global_size or v.shape[0] != space.global_size:
        raise ValueError("u and v must have one entry per DOF")
    return BilinearForm.apply(points, u, v, assembler, space)
