"""torch.autograd bindings of the k-space and total forces and their adjoints.

`KspaceForce.apply(pos, q, solver)` returns the k-space force (3, n) of an
initialised Solver for device positions pos (3, n) and charges q (n), FP64
CUDA tensors; the backward pass calls the library's hand-written adjoint
(p3m_hip_kspace_force_vjp_device) through ctypes.  Nothing of the forward
pass is stored except the inputs: the adjoint recomputes the field meshes.

The force returned is f_k plus, for methods 2 and 7, the self-force term
that Kspace_force subtracts from f (p3m-ad-self-forces.c:114-139).

The library works on its own HIP stream: each call waits for torch's current
stream before enqueueing and for the library's stream before returning.
"""
import torch


def _check(pos, q, solver):
    n = q.numel()
    if pos.shape != (3, n):
        raise ValueError("positions must have shape (3, n), got %s" % (tuple(pos.shape),))
    if n != solver.system.n:
        raise ValueError("the solver was initialised for %d particles, not %d" % (solver.system.n, n))
    for t in (pos, q):
        if t.dtype != torch.float64 or not t.is_cuda:
            raise TypeError("positions and charges must be float64 CUDA tensors")


class KspaceForce(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, q, solver):
        _check(pos, q, solver)
        n = q.numel()
        pos = pos.contiguous()
        q = q.contiguous()
        L = float(solver.system.c.length)
        fk = torch.zeros(3 * n, dtype=torch.float64, device=pos.device)
        f = torch.zeros_like(fk)
        torch.cuda.current_stream().synchronize()
        solver.kspace_force_device(n, L, pos[0].data_ptr(), pos[1].data_ptr(), pos[2].data_ptr(),
                                   q.data_ptr(), fk.data_ptr(), f.data_ptr())
        solver.synchronize()
        ctx.save_for_backward(pos, q)
        ctx.solver = solver
        return (fk + f).view(3, n)

    @staticmethod
    def backward(ctx, gF):
        pos, q = ctx.saved_tensors
        solver = ctx.solver
        n = q.numel()
        g = gF.to(torch.float64).contiguous().view(3 * n)
        dpos = torch.zeros(3 * n, dtype=torch.float64, device=pos.device)
        dq = torch.zeros(n, dtype=torch.float64, device=pos.device)
        torch.cuda.current_stream().synchronize()
        solver.kspace_force_vjp_device(n, float(solver.system.c.length), pos[0].data_ptr(),
                                       pos[1].data_ptr(), pos[2].data_ptr(), q.data_ptr(),
                                       g.data_ptr(), dpos.data_ptr(), dq.data_ptr())
        solver.synchronize()
        return dpos.view(3, n), dq, None


def kspace_force(pos, q, solver):
    """Differentiable k-space force (3, n) of `solver` at pos (3, n), q (n)."""
    return KspaceForce.apply(pos, q, solver)


class TotalForce(torch.autograd.Function):
    """Calculate_forces (common.c:268-327): f = f_k + f_r (ad / ad-r with the
    self-force term), differentiable in positions and charges; the backward
    pass adds the k-space adjoint and the real-space adjoint
    (p3m_hip_realteil_vjp_device)."""

    @staticmethod
    def forward(ctx, pos, q, solver, rcut, alpha):
        _check(pos, q, solver)
        n = q.numel()
        pos = pos.contiguous()
        q = q.contiguous()
        L = float(solver.system.c.length)
        f = torch.empty(3 * n, dtype=torch.float64, device=pos.device)
        fk = torch.empty_like(f)
        fr = torch.empty_like(f)
        torch.cuda.current_stream().synchronize()
        solver.calculate_forces_device(n, L, pos[0].data_ptr(), pos[1].data_ptr(), pos[2].data_ptr(),
                                       q.data_ptr(), float(rcut), float(alpha), f.data_ptr(),
                                       fk.data_ptr(), fr.data_ptr(), None)
        solver.synchronize()
        ctx.save_for_backward(pos, q)
        ctx.solver, ctx.rcut, ctx.alpha = solver, float(rcut), float(alpha)
        return f.view(3, n)

    @staticmethod
    def backward(ctx, gF):
        pos, q = ctx.saved_tensors
        solver = ctx.solver
        n = q.numel()
        L = float(solver.system.c.length)
        g = gF.to(torch.float64).contiguous().view(3 * n)
        dpos = torch.zeros(3 * n, dtype=torch.float64, device=pos.device)
        dq = torch.zeros(n, dtype=torch.float64, device=pos.device)
        ptrs = (pos[0].data_ptr(), pos[1].data_ptr(), pos[2].data_ptr(), q.data_ptr())
        torch.cuda.current_stream().synchronize()
        solver.kspace_force_vjp_device(n, L, *ptrs, g.data_ptr(), dpos.data_ptr(), dq.data_ptr())
        if ctx.rcut != 0.0:
            solver.realteil_vjp_device(n, L, *ptrs, ctx.rcut, ctx.alpha, g.data_ptr(),
                                       dpos.data_ptr(), dq.data_ptr())
        solver.synchronize()
        return dpos.view(3, n), dq, None, None, None


def total_force(pos, q, solver, rcut, alpha):
    """Differentiable total force (3, n): k space of `solver` plus real space
    at (rcut, alpha); rcut = 0 leaves the real-space part out."""
    return TotalForce.apply(pos, q, solver, rcut, alpha)
