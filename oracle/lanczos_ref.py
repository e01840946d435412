"""FP64 Lanczos restatement used as the eigenvalue oracle (test infrastructure).

Follows the reference's `lanczos_p` loop (src/blanczosmain.f90:1582-1753):
uniform pivot 1/sqrt(D) (src/blanczoslib.f90:825-831); w = H v_i;
alpha_i = <v_i|w>, w -= alpha_i v_i (br_remove_prev_overlap,
src/breorthog.f90:683-711); classical Gram-Schmidt against every stored vector,
all dot products first and then one subtraction (br_orthogonalize, :831-891);
beta_i = ||w||, v_{i+1} = w / beta_i (br_normalize, :758-783); Ritz values from
the tridiagonal matrix (find_lanczos_eigenvalues, src/blanczoslib.f90:143-376).
Everything is FP64 here (the reference keeps vectors in FP32 by default).
"""
import numpy as np


def lanczos(apply, dim, niter, nkeep):
    niter = min(niter, dim)
    V = np.zeros((niter + 1, dim))
    v = np.full(dim, 1.0 / np.sqrt(dim))
    V[0] = v
    alpha = np.zeros(niter)
    beta = np.zeros(niter)
    for i in range(niter):
        w = apply(V[i])
        a = float(V[i] @ w)
        w = w - a * V[i]
        d = V[: i + 1] @ w
        w = w - d @ V[: i + 1]
        b = float(np.linalg.norm(w))
        alpha[i] = a
        beta[i] = b
        if i + 1 < niter + 1:
            V[i + 1] = w / b if b > 0 else 0.0
    T = np.diag(alpha) + np.diag(beta[:-1], 1) + np.diag(beta[:-1], -1)
    e, y = np.linalg.eigh(T)
    vecs = y[:, :nkeep].T @ V[:niter]
    return alpha, beta, e[:nkeep], vecs


def convergence_stop(alpha, beta, nkeep, ncheck_ex=5, tol=1e-3, test=0, niter=None):
    """Iteration at which a plain Lanczos run with the reference's default
    convergence stops (the rule of lanczos_p, src/blanczosmain.f90:1469-1512,
    :1690-1724, and convergence_test, src/blanczoslib.f90:2660-2800): every
    itskipeig-th iteration (10, or int(sqrt(nkeep)) if larger) beyond ncheck =
    nkeep + ncheck_ex, the lowest ncheck eigenvalues of the tridiagonal so far
    are compared with those of the previous such check; test 0 stops when
    sum |dE| / sqrt(ncheck + 1) < tol, test 1 when max |dE| < tol.  Returns the
    stopping iteration (1-based): niter when the test never passes before it,
    None when the coefficients end before either."""
    niter = len(alpha) if niter is None else niter
    itskip = max(10, int(np.sqrt(nkeep)))
    ncheck = nkeep + ncheck_ex
    ntest = ncheck if nkeep <= 100 else int(np.sqrt(ncheck) + 1)
    eold = None
    for it in range(1, len(alpha) + 1):
        if it >= niter:
            return niter
        if it % itskip or it <= ncheck:
            continue
        T = np.diag(alpha[:it]) + np.diag(beta[:it - 1], 1) + np.diag(beta[:it - 1], -1)
        e = np.linalg.eigvalsh(T)[:ncheck]
        if eold is not None:
            d = np.abs(eold[:ntest] - e[:ntest])
            if (d.max() if test == 1 else d.sum() / np.sqrt(ntest + 1)) < tol:
                return it
        eold = e
    return None
