"""Lanczos breakdown handling on the GPU (bs_lanczos / bs_strength), following
lanczos_p (src/blanczosmain.f90:1634-1658): a beta below 1e-4 of the previous
one (|H v| on the first step) restarts from a seeded random vector
orthonormalized against the history (src/breorthog.f90:1471-1482); a beta
below 0.1 of the previous one triggers the orthogonality test and a second
Gram-Schmidt pass (src/breorthog.f90:977-1045).

The reference tests only from its second iteration on, so a pivot that is an
exact eigenvector is normalized noise there; here the first step is tested
against |H v|.  Everything is checked against the oracle's dense spectrum."""
import numpy as np
import pytest

from golden_cases import paths, small_fixture
from oracle_api import Oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _setup(name, casedir):
    from bigstickpublick_amd import System
    case, d, _ = small_fixture(name, casedir)
    s = System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    ev, U = np.linalg.eigh(o.dense())
    return s, ev, U


def _near_spectrum(ritz, ev, tol):
    """Distance of every value in ritz to the nearest eigenvalue."""
    return np.abs(ritz[:, None] - ev[None, :]).min(axis=1).max() < tol


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_strength_from_exact_eigenvector(dtype, casedir):
    """bs_strength from an eigenvector of bs_exact (sd 2p2n, D = 640): the
    first beta is rounding noise, the run restarts from a random vector, and
    every output is finite; the whole strength sits on that eigenvalue and
    the later Ritz values are eigenvalues of H."""
    s, ev, U = _setup("sd2p2n", casedir)
    ex = s.exact(6, dtype="f64", want_vectors=True)
    assert np.abs(ex.evals - ev[:6]).max() < 1e-10
    k = 2
    piv = ex.evecs[k].clone() if dtype == "f64" else ex.evecs[k].float()
    r = s.strength(piv, nkeep=5, niter=200, polish=True)
    rr, sp, _, _ = s.lanczos_stats()
    assert rr >= 1
    for a in (r.ritz, r.strength, r.evals, r.alpha, r.beta, r.resid):
        assert np.all(np.isfinite(a))
    assert abs(r.pivot_norm - 1.0) < 1e-6
    tol_e = 1e-8 if dtype == "f64" else 2e-4
    top = int(np.argmax(r.strength))
    assert abs(r.ritz[top] - ev[k]) < tol_e
    assert r.strength[top] > 1.0 - (1e-10 if dtype == "f64" else 1e-5)
    assert abs(r.strength.sum() - 1.0) < 1e-6
    # the first beta is the breakdown, kept small as the reference keeps it
    assert r.beta[0] < 1e-4 * abs(r.alpha[0])
    # 200 steps on D = 640: the low end of the spectrum has converged, so
    # the lowest Ritz values are the lowest eigenvalues, eigenvalue k included
    assert np.abs(np.sort(r.ritz)[:5] - ev[:5]).max() < (1e-8 if dtype == "f64" else 2e-4)
    assert np.abs(r.evals - ev[:5]).max() < (1e-8 if dtype == "f64" else 1e-5)


def test_pivot_in_two_dimensional_invariant_subspace(casedir):
    """Pivot (u_0 + u_3)/sqrt(2): the Krylov space closes after two vectors,
    beta_1 collapses against beta_0, the run restarts once and still returns
    the dense spectrum's lowest eigenvalues; the strength is 1/2 on E_0 and
    E_3."""
    s, ev, U = _setup("sd2p2n", casedir)
    p = (U[:, 0] + U[:, 3]) / np.sqrt(2.0)
    r = s.strength(torch.tensor(p, device="cuda"), nkeep=5, niter=200, polish=True)
    rr = s.lanczos_stats()[0]
    assert rr >= 1
    assert np.all(np.isfinite(r.ritz)) and np.all(np.isfinite(r.strength))
    assert r.beta[1] < 1e-4 * r.beta[0]
    for k in (0, 3):
        j = int(np.argmin(np.abs(r.ritz - ev[k])))
        assert abs(r.ritz[j] - ev[k]) < 1e-9
        assert abs(r.strength[j] - 0.5) < 1e-9
    assert np.abs(r.evals - ev[:5]).max() < 1e-8
    assert _near_spectrum(np.sort(r.ritz)[:6], ev, 1e-6)


def test_plain_run_without_breakdown_is_unchanged(casedir):
    """sd 1p2n, 2Jz = 1 (D = 128; the uniform pivot overlaps every low
    eigenvector): no step breaks down, no restart is taken, and the FP64
    coefficients and Ritz values equal the oracle's FP64 recurrence."""
    from lanczos_ref import lanczos
    from golden_cases import paths as _paths
    from bigstickpublick_amd import System
    case, d, _ = small_fixture("sd1p2n_m1", casedir)
    s = System(*_paths(case, d), case.Z, case.N, case.jz2, upload=True)
    o = Oracle(*_paths(case, d), case.Z, case.N, case.jz2)
    a, b, e, _ = lanczos(o.apply, o.dim, 60, 5)
    r = s.lanczos(nkeep=5, niter=60, dtype="f64", polish=False)
    rr = s.lanczos_stats()[0]
    assert rr == 0
    assert np.abs(r.alpha[:40] - a[:40]).max() < 1e-9
    assert np.abs(r.evals - e).max() < 1e-8


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_thick_restart_with_krylov_space_as_large_as_basis(dtype, casedir):
    """sd 1p1n (D = 28) with restart length 40 >= D: one cycle spans the
    basis; the last beta is zero, no restart is taken, and the eigenvalues are
    the dense ones."""
    s, ev, U = _setup("sd1p1n", casedir)
    r = s.lanczos(nkeep=4, niter=40, thick=10, max_restarts=5, dtype=dtype, tol=1e-12, polish=True)
    assert np.all(np.isfinite(r.evals)) and np.all(np.isfinite(r.resid))
    assert r.matvecs <= 28 + 4
    assert np.abs(r.evals - ev[:4]).max() < (1e-10 if dtype == "f64" else 2e-5)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_breakdown_full_space(dtype, tmp_path):
    """Three Hamiltonians on sd 1p1n (D = 28) and on sd 1p2n, 2Jz = 1
    (D = 128), in one batched run each, with Krylov spaces as large as the
    basis: the uniform pivot's invariant subspace closes before the basis is
    spanned, the device-side breakdown handling restarts those members, every
    output is finite and the Ritz values equal each member's dense
    spectrum."""
    from bigstickpublick_amd import Batch
    from bigstickpublick_amd.inputs import small_case, write_case
    for (Z, N, jz2, niter) in ((1, 1, 0, 40), (1, 2, 1, 140)):
        cases = []
        for m in range(3):
            c = small_case("sd", Z, N, jz2, seed=500 + m)
            c.name = "bk%d_%d" % (Z, m)
            write_case(c, str(tmp_path))
            cases.append(c)
        ints = [str(tmp_path / (c.name + ".int")) for c in cases]
        b = Batch(str(tmp_path / "sd.sps"), ints, Z, N, jz2)
        r = b.lanczos(nkeep=4, niter=niter, dtype=dtype, polish=True)
        assert np.all(np.isfinite(r.evals)) and np.all(np.isfinite(r.alpha)) and np.all(np.isfinite(r.beta))
        for k, c in enumerate(cases):
            o = Oracle(str(tmp_path / "sd.sps"), ints[k], Z, N, jz2)
            ev = np.linalg.eigvalsh(o.dense())
            assert np.abs(r.evals[k] - ev[:4]).max() < (1e-9 if dtype == "f64" else 2e-5), (Z, k)
        b.close()
