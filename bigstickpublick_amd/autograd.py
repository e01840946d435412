"""torch.autograd.Function over the engine's Lanczos and adjoint passes.

Forward: interaction parameters theta = (SPEs, TBMEs in .int order) -> lowest
eigenvalues, by writing the iso .int file (values rounded to FP32, as the
reference reads them) and running the batched Lanczos (bs_batch_lanczos) on
it.  The basis and the jump tables are built once per model space, particle
numbers, 2Jz, number of Hamiltonians and list of two-body elements, and kept
in a small cache: a later call only replaces the interaction values
(bs_batch_set_interactions), so an optimisation loop pays for the tables once.
Backward: dE_k/dtheta = <psi_k| dH/dtheta |psi_k> (Hellmann-Feynman) from the
adjoint density pass, one per eigenvector (bs_batch_gradient).

eigenenergies: one Hamiltonian, theta of shape (P,) -> energies (nkeep,)
ensemble_energies: K Hamiltonians on one basis, thetas (K, P) -> (K, nkeep)
"""
import collections
import os
import tempfile

import numpy as np
import torch

from .engine import Batch
from .inputs import Interaction, sps_text

_CACHE_SIZE = 4
_cache = collections.OrderedDict()  # key -> Batch, least recently used first


def clear_cache():
    """Release the cached batches (their basis and jump tables)."""
    while _cache:
        _, b = _cache.popitem(last=False)
        b.close()


def _key(template, Z, N, jz2, K):
    return (tuple(template.orbits), int(Z), int(N), int(jz2), int(K), tuple(l[:6] for l in template.tbmes))


def _batch(template, Z, N, jz2, thetas):
    """The cached batch for this space and line list, holding the values of
    thetas (one row per Hamiltonian)."""
    K = thetas.shape[0]
    norb = len(template.orbits)
    key = _key(template, Z, N, jz2, K)
    with tempfile.TemporaryDirectory(prefix="bsgrad_") as d:
        files = []
        for k in range(K):
            it = Interaction(template.orbits, tuple(float(v) for v in thetas[k, :norb]),
                             [l[:6] + (float(v),) for l, v in zip(template.tbmes, thetas[k, norb:])])
            f = os.path.join(d, "h%d.int" % k)
            with open(f, "w") as fh:
                fh.write(it.int_text())
            files.append(f)
        b = _cache.get(key)
        if b is None:
            sps = os.path.join(d, "space.sps")
            with open(sps, "w") as fh:
                fh.write(sps_text(template.orbits))
            b = Batch(sps, files, Z, N, jz2)
            _cache[key] = b
            while len(_cache) > _CACHE_SIZE:
                _cache.popitem(last=False)[1].close()
        else:
            b.set_interactions(files)
            _cache.move_to_end(key)
    return b


class EnsembleEnergies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, thetas, template, Z, N, jz2, nkeep, niter, thick, dtype):
        th = thetas.detach().cpu().double().numpy()
        if th.ndim != 2 or th.shape[1] != len(template.orbits) + len(template.tbmes):
            raise ValueError("thetas must have shape (K, %d)" % (len(template.orbits) + len(template.tbmes)))
        b = _batch(template, Z, N, jz2, th)
        r = b.lanczos(nkeep=nkeep, niter=niter, thick=thick, max_restarts=50 if thick else 0, dtype=dtype,
                      tol=1e-10 if thick else 0.0, polish=True, want_vectors=True)
        ctx.batch = b
        ctx.vecs = r.evecs  # (nkeep, dim, K)
        ctx.nkeep = nkeep
        return torch.tensor(r.evals, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        # the gradient needs only the eigenvectors and the (value-independent)
        # Jacobian, so a later forward that changed the batch's values does
        # not affect it
        b = ctx.batch
        G = np.stack([b.gradient(ctx.vecs[n]) for n in range(ctx.nkeep)])  # (nkeep, K, P)
        g = np.einsum("kn,nkp->kp", gE.detach().cpu().double().numpy(), G)
        return torch.tensor(g, dtype=torch.float64), None, None, None, None, None, None, None, None


def ensemble_energies(thetas, template, Z, N, jz2=0, nkeep=5, niter=150, thick=0, dtype="f64"):
    """Differentiable lowest nkeep energies of K Hamiltonians on one basis:
    thetas (K, P), each row the SPEs then TBMEs in the order of `template`;
    returns (K, nkeep).  Plain Lanczos (thick = 0, niter <= 256) or thick
    restart (thick <= 16, tested to 1e-10 MeV over restarts)."""
    return EnsembleEnergies.apply(thetas, template, Z, N, jz2, nkeep, niter, thick, dtype)


def eigenenergies(theta, template, Z, N, jz2=0, nkeep=5, niter=150, thick=0, dtype="f64"):
    """Differentiable lowest nkeep energies of the Hamiltonian with parameters
    theta (SPEs then TBMEs, in the order of `template`)."""
    return EnsembleEnergies.apply(theta.reshape(1, -1), template, Z, N, jz2, nkeep, niter, thick, dtype)[0]


def parameters(interaction):
    """theta = (SPEs, TBME values) of an Interaction, as a float64 tensor."""
    return torch.tensor(list(interaction.spe) + [t[6] for t in interaction.tbmes], dtype=torch.float64)
