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
    sum |dE| / sqrt(ncheck + 1) < tol, test 1 when max |dE| < tol.  Tests 2
    and 3 (altconverge, src/blanczoslib.f90:2741-2756) compare the check's
    lowest ncheck eigenvectors of the tridiagonal with the previous check's
    instead: vdot_j = sum over components k < it - 1 of the two (the earlier
    vector zero beyond its length: find_lanczos_eigenvalues zeroes the array
    before each solve, src/blanczoslib.f90:196), test 2 stops when
    sqrt(sum_j (|vdot_j| - 1)^2) / ncheck < tol, test 3 when max_j
    ||vdot_j| - 1| < tol.  A plain run solves for the lowest nkeep
    eigenvectors only (DSYEVR over indices 1..nkeep, src/blanczoslib.f90:
    230-234, with use_switch_esolver on by default), so the ncheck_ex extra
    columns are zero, their vdot is 0, and with ncheck_ex > 0 tests 2 and 3
    never pass.  Returns the stopping iteration (1-based): niter when the test
    never passes before it, None when the coefficients end before either."""
    niter = len(alpha) if niter is None else niter
    itskip = max(10, int(np.sqrt(nkeep)))
    ncheck = nkeep + ncheck_ex
    ntest = ncheck if nkeep <= 100 else int(np.sqrt(ncheck) + 1)
    eold = None
    yold = None
    for it in range(1, len(alpha) + 1):
        if it >= niter:
            return niter
        if it % itskip or it <= ncheck:
            continue
        T = np.diag(alpha[:it]) + np.diag(beta[:it - 1], 1) + np.diag(beta[:it - 1], -1)
        if test in (2, 3):
            e, V = np.linalg.eigh(T)
            y = V[:, :ncheck].copy()
            y[:, nkeep:] = 0.0
            if yold is not None:
                n = min(it - 1, yold.shape[0])
                vd = np.abs(np.einsum("kj,kj->j", y[:n], yold[:n])) - 1.0
                v = np.sqrt((vd ** 2).sum()) / ncheck if test == 2 else np.abs(vd).max()
                if v < tol:
                    return it
            yold = y
            continue
        e = np.linalg.eigvalsh(T)[:ncheck]
        if eold is not None:
            d = np.abs(eold[:ntest] - e[:ntest])
            if (d.max() if test == 1 else d.sum() / np.sqrt(ntest + 1)) < tol:
                return it
        eold = e
    return None


def target_order(e, Y, nkeep, etarget):
    """Ritz pairs as the reference's option tx orders them (txchoice 1 of
    find_lanczos_eigenvalues, src/blanczoslib.f90:277-327): a selection sort
    by |E - etarget| (first minimum kept on ties), then a selection sort of
    the first nkeep by energy.  e ascending; Y columns are the pairs."""
    e = e.copy()
    Y = Y.copy()
    d = np.abs(e - etarget)
    n = len(e)
    for i in range(n - 1):
        k = i + int(np.argmin(d[i:]))
        if k != i:
            d[[i, k]] = d[[k, i]]
            e[[i, k]] = e[[k, i]]
            Y[:, [i, k]] = Y[:, [k, i]]
    m = min(nkeep, n)
    for i in range(m - 1):
        k = i + int(np.argmin(e[i:m]))
        if k != i:
            e[[i, k]] = e[[k, i]]
            Y[:, [i, k]] = Y[:, [k, i]]
    return e, Y


def thick_lanczos(apply, dim, nkeep, niter, nthick, max_restarts, tol=0.0, test=0, ncheck_ex=0,
                  etarget=None, target_test="reference"):
    """Thick-restart Lanczos as the reference runs it (options td, tf, tc,
    tx: lanczos_p, src/blanczosmain.f90:1582-1752; thick_restart_sub_p,
    src/blanczoslib.f90:1656-1799; convergence_test, :2660-2800), in FP64,
    from the uniform pivot, without the breakdown restarts.

    iter counts the vectors of the current projected matrix.  Each step
    applies H to v_iter, takes alpha, orthogonalizes against every earlier
    vector and normalizes.  The Ritz values are taken when iter is a
    multiple of itskipeig = max(10, int(sqrt(nkeep))) beyond max(ncheck,
    nkeep), and at iter = niter - 1; at every multiple beyond ncheck (ncheck
    = nkeep + ncheck_ex, nkeep for a targeted run) the test compares them
    with the previous iteration's: test 0 sum |dE| / sqrt(ntest + 1) times
    (1 + restarts), test 1 max |dE|, over the first ntest values; tests 2 and
    3 compare the first ncheck Ritz vectors with the previous solve's instead
    (as convergence_stop does; every vector exists in a thick run).  At iter =
    niter - 1 a run that has not converged restarts -- after max_restarts + 1
    restarts it stops instead: the matrix on the first niter - 2 vectors is
    diagonalized, its nthick leading Ritz vectors replace v_1 .. v_nthick,
    couple to v_{niter-1} (the next v_{nthick+1}) by beta_{niter-2}
    Y[niter-2, k], and iter resumes at nthick.  The projected matrix holds the
    kept Ritz values on its diagonal with those couplings in column nthick+1,
    then the tridiagonal part.  With etarget every Ritz solve is ordered by
    target_order; the reference's tx tests no energies (its count of tested
    energies is never set on that path, src/blanczosmain.f90:1497-1499),
    target_test="kept" tests the nkeep kept ones.

    Returns a dict: evals (nkeep), alpha and beta of the final matrix (the
    reference's .lcoef: the kept values and couplings, then the last
    cycle's steps), steps (its size), restarts, matvecs."""
    itskip = max(10, int(np.sqrt(nkeep)))
    ncheck = min(dim, nkeep + (0 if etarget is not None else ncheck_ex))
    ntest = 0 if (etarget is not None and target_test == "reference") else ncheck
    V = np.zeros((niter + 1, dim))
    V[0] = 1.0 / np.sqrt(dim)
    alpha = np.zeros(niter + 1)
    beta = np.zeros(niter + 1)
    kstart = 0
    restarts = 0
    matvecs = 0

    def ritz(m):
        A = np.diag(alpha[:m])
        for i in range(kstart, m - 1):
            A[i, i + 1] = A[i + 1, i] = beta[i]
        if kstart < m:
            A[:kstart, kstart] = beta[:kstart]
            A[kstart, :kstart] = beta[:kstart]
        e, Y = np.linalg.eigh(A)
        if etarget is not None:
            e, Y = target_order(e, Y, nkeep, etarget)
        return e, Y

    def change(e0, e1):
        d = np.abs(e0[:ntest] - e1[:ntest])
        if ntest == 0:
            return 0.0
        return d.max() if test == 1 else d.sum() / np.sqrt(ntest + 1) * (1 + restarts)

    def vchange(Yn, Yo, it):
        # tests 2, 3 (src/blanczoslib.f90:2741-2756): the first ncheck Ritz
        # vectors against the previous solve's, components below it - 1, the
        # earlier ones zero beyond their length (none before the first solve)
        vd = np.zeros(ncheck)
        if Yo is not None:
            n = min(it - 1, Yn.shape[0], Yo.shape[0])
            c = min(ncheck, Yn.shape[1], Yo.shape[1])
            vd[:c] = np.einsum("kj,kj->j", Yn[:n, :c], Yo[:n, :c])
        d = np.abs(vd) - 1.0
        return np.abs(d).max() if test == 3 else np.sqrt((d ** 2).sum()) / ncheck

    ecur = np.zeros(ncheck)
    eold = np.zeros(ncheck)
    ylast = None  # Ritz vectors of the last solve
    it = 0
    finished = False
    while not finished:
        it += 1
        i = it - 1
        w = apply(V[i])
        matvecs += 1
        a = float(V[i] @ w)
        w = w - a * V[i]
        d = V[:i] @ w
        w = w - d @ V[:i]
        b = float(np.linalg.norm(w))
        alpha[i] = a
        beta[i] = b
        V[i + 1] = w / b
        fly = it % itskip == 0
        yprev = ylast
        if (fly and it > max(ncheck, nkeep)) or it + 1 == niter:
            e, ylast = ritz(it)
            ecur = np.array([e[j] if j < it else 0.0 for j in range(ncheck)])
        if it == niter:
            finished = True
        elif tol > 0.0:
            if it > ncheck and fly:
                if test in (2, 3):
                    finished = vchange(ylast, yprev, it) < tol
                elif change(eold, ecur) < tol:
                    finished = True
            eold = ecur.copy()
        if finished or it + 1 != niter:
            continue
        if restarts > max_restarts:
            break
        m = it - 1
        e, Y = ritz(m)
        ylast = Y
        ecur = np.array([e[j] if j < m else 0.0 for j in range(ncheck)])
        keep = Y[:, :nthick].T @ V[:m]
        V[nthick] = V[m].copy()
        V[:nthick] = keep
        beta[:nthick] = beta[m - 1] * Y[m - 1, :nthick]
        alpha[:nthick] = e[:nthick]
        kstart = nthick
        restarts += 1
        it = nthick
    e, Y = ritz(it)
    return {"evals": e[:nkeep], "alpha": alpha[:it].copy(), "beta": beta[:it].copy(), "steps": it,
            "restarts": restarts, "matvecs": matvecs}
