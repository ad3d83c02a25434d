"""Delta-mu as a differentiable function of the solute parameters.

`solvation_free_energy(xyz, q, sig, eps, n, box, solvent, **opts)` runs one
3D-RISM solve on the GPU (rism_create / rism_build_potentials / rism_solve /
rism_thermo) and returns SFE_SC (J/mol) as a 0-d float64 tensor.  Its
backward pass is the adjoint of the converged fixed point
(rism_adjoint_grad), so gradients flow to any of xyz [n][3] (Angstrom),
q [n] (e), sig [n] (Angstrom) and eps [n] (J/mol) that require them.
"""
import numpy as np
import torch

from . import FP64, Options, Solute, Solver


class SolvationFreeEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, q, sig, eps, n, box, solvent, opts, rtol):
        su = Solute(sig.detach().cpu().double().numpy(), eps.detach().cpu().double().numpy(),
                    q.detach().cpu().double().numpy(), xyz.detach().cpu().double().numpy().reshape(-1, 3))
        s = Solver(n, box, solvent, su, opts)
        try:
            s.build_potentials()
            st, steps, rms = s.solve()
            if st != 0:
                raise RuntimeError(f"solvation_free_energy: solve ended with status {st} (rms {rms:.3e})")
            x = s.thermo()
            g = s.adjoint_grad(rtol) if any(ctx.needs_input_grad[:4]) else None
        finally:
            s.close()
        dev, dt = xyz.device, xyz.dtype
        if g is not None:
            ctx.save_for_backward(*(torch.as_tensor(g[k], dtype=dt, device=dev) for k in ("xyz", "q", "sig", "eps")))
        return torch.tensor(x["SFE_SC"], dtype=dt, device=dev)

    @staticmethod
    def backward(ctx, gout):
        dxyz, dq, dsig, deps = ctx.saved_tensors
        return gout * dxyz, gout * dq, gout * dsig, gout * deps, None, None, None, None, None


def solvation_free_energy(xyz, q, sig, eps, n, box, solvent, rtol=1e-10, **opts):
    """SFE_SC (J/mol) of the solute (xyz, q, sig, eps) in `solvent` on an
    n[0] x n[1] x n[2] grid of `box` Angstrom; differentiable in the four
    solute tensors.  opts are rism3d.Options fields (closure, tol, ...)."""
    o = Options(**{"precision": FP64, **opts})
    return SolvationFreeEnergy.apply(xyz, q, sig, eps, tuple(n), tuple(box), solvent, o, rtol)
