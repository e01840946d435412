"""torch.autograd.Function over the engine's Lanczos and adjoint passes.

Forward: interaction parameters theta = (SPEs, TBMEs in .int order) -> lowest
eigenvalues E_k, by writing the iso .int file (values rounded to FP32, as the
reference reads them), building the tables and running bs_lanczos.
Backward: dE_k/dtheta = <psi_k| dH/dtheta |psi_k> (Hellmann-Feynman) from
bs_gradient, one adjoint pass per eigenvector.
"""
import os
import tempfile

import numpy as np
import torch

from .engine import System
from .inputs import Interaction, sps_text


class EigenEnergies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, template, Z, N, jz2, nkeep, niter, thick, dtype):
        th = theta.detach().cpu().double().numpy()
        norb = len(template.orbits)
        it = Interaction(template.orbits, tuple(float(v) for v in th[:norb]),
                         [l[:6] + (float(v),) for l, v in zip(template.tbmes, th[norb:])])
        # the files are read once, by the host table build; the directory
        # goes away as soon as the system holds its tables
        with tempfile.TemporaryDirectory(prefix="bsgrad_") as d:
            sps, intf = os.path.join(d, "space.sps"), os.path.join(d, "h.int")
            with open(sps, "w") as f:
                f.write(sps_text(it.orbits))
            with open(intf, "w") as f:
                f.write(it.int_text())
            s = System(sps, intf, Z, N, jz2, upload=True)
        r = s.lanczos(nkeep=nkeep, niter=niter, thick=thick, max_restarts=50 if thick else 0,
                      dtype=dtype, tol=1e-10 if thick else 0.0, polish=True, want_vectors=True)
        ctx.system = s
        ctx.vecs = r.evecs
        ctx.nkeep = nkeep
        return torch.tensor(r.evals, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        s = ctx.system
        G = np.stack([s.gradient(ctx.vecs[k]) for k in range(ctx.nkeep)])
        g = torch.tensor(gE.detach().cpu().double().numpy() @ G, dtype=torch.float64)
        return g, None, None, None, None, None, None, None, None


def eigenenergies(theta, template, Z, N, jz2=0, nkeep=5, niter=150, thick=0, dtype="f64"):
    """Differentiable lowest nkeep energies of the Hamiltonian with parameters
    theta (SPEs then TBMEs, in the order of `template`)."""
    return EigenEnergies.apply(theta, template, Z, N, jz2, nkeep, niter, thick, dtype)


def parameters(interaction):
    """theta = (SPEs, TBME values) of an Interaction, as a float64 tensor."""
    return torch.tensor(list(interaction.spe) + [t[6] for t in interaction.tbmes], dtype=torch.float64)
