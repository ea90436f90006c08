"""mu of a solute as a differentiable torch function.

`solute_mu(params, names, solvent, settings, chi)` returns the excess
chemical potential of the solute whose site parameters are the rows of
`params` (n, 6): x, y, z [A], charge, sigma [A], epsilon [kcal/mol].  The
forward pass is rism3d_solute_solve on the GPU; its backward pass is the
adjoint gradient the same call returns (rism3d_result.dmu_dsite: the
transposed linearised closure + RISM equation solved at the converged t,
reusing the residual's FFT kernels, no iterates stored), so one forward
and backward cost one solve plus one linear solve.  The gradient is exact
for the discretised problem, on the tau-renormalised path (the default)
and the Ng path (`no-renorm`); with `precision` "mixed" the linear solve
runs by FP32 defect correction as the forward solve does.  Tensors live on
the CPU: the library owns the device memory.
"""
import numpy as np
import torch

from . import rism3d as R

COLUMNS = ("x", "y", "z", "charge", "sigma", "epsilon")


def sites_from_params(names, params):
    """(n, 6) array in COLUMNS order -> the (name, xyz, sigma, epsilon,
    charge) site tuples of bgy3d_amd.rism3d."""
    p = np.asarray(params, dtype=np.float64)
    if p.ndim != 2 or p.shape[1] != 6 or p.shape[0] != len(names):
        raise ValueError(f"params must be ({len(names)}, 6), got {p.shape}")
    return [(nm, (p[a, 0], p[a, 1], p[a, 2]), p[a, 4], p[a, 5], p[a, 3])
            for a, nm in enumerate(names)]


def params_from_sites(sites):
    """Inverse of sites_from_params: (names, (n, 6) float64 array)."""
    names = [s[0] for s in sites]
    p = np.array([[s[1][0], s[1][1], s[1][2], s[4], s[2], s[3]] for s in sites],
                 dtype=np.float64)
    return names, p


class SoluteMu(torch.autograd.Function):
    """Forward: mu (0-d tensor).  Backward: grad_mu * d mu / d params."""

    @staticmethod
    def forward(ctx, params, names, solvent, settings, chi):
        want = bool(ctx.needs_input_grad[0])
        sites = sites_from_params(names, params.detach().cpu().numpy())
        out = R.hnc3d_run_solute(sites, solvent, settings, chi=chi, adjoint=want)
        if out["status"] != "OK":
            raise R.RismError(getattr(R, out["status"]), "solute_mu: solve did not converge")
        if want:
            ctx.save_for_backward(torch.as_tensor(out["free-energy-dsite"], dtype=torch.float64))
        return params.new_tensor(out["free-energy"])

    @staticmethod
    def backward(ctx, grad_mu):
        (d,) = ctx.saved_tensors
        return grad_mu * d.to(grad_mu.device, grad_mu.dtype), None, None, None, None


def solute_mu(params, names, solvent, settings, chi=None):
    """mu(params) with autograd; see the module docstring."""
    return SoluteMu.apply(params, list(names), solvent, dict(settings), chi)
