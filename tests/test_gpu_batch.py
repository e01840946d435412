"""Batched Hamiltonians on one basis (the sd 4p4n ensemble) and the block
apply, through the C ABI, against the oracle and the single-Hamiltonian path."""
import os

import numpy as np
import pytest

from bigstickpublick_amd.inputs import config_case, f32, write_case
from golden_cases import config_fixture, ensemble_fixture, paths
from lanczos_ref import lanczos
from oracle_api import Oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _batch(casedir, K):
    from bigstickpublick_amd import Batch
    cases, d, fx = ensemble_fixture(casedir, K)
    b = Batch(os.path.join(d, "sd.sps"), [paths(c, d)[1] for c in cases], 4, 4, 0)
    return cases, d, fx, b


def test_batch_apply_matches_oracle(casedir):
    cases, d, fx, b = _batch(casedir, 8)
    assert (b.K, b.dim) == (8, 28503)
    X = np.random.default_rng(5).standard_normal((b.dim, b.K))
    Y64 = b.apply(torch.tensor(X, device="cuda")).cpu().numpy()
    Y32 = b.apply(torch.tensor(X, device="cuda", dtype=torch.float32)).cpu().numpy()
    for k, c in enumerate(cases):
        want = Oracle(*paths(c, d), c.Z, c.N, c.jz2).apply(np.ascontiguousarray(X[:, k]))
        assert _rel(Y64[:, k], want) < 1e-12, k
        assert _rel(Y32[:, k], want) < 2e-5, k


def test_block_apply_matches_single(casedir):
    """Same H on K interleaved vectors (the reference's block apply) equals K
    single applies."""
    from bigstickpublick_amd import System
    case, d, fx = config_fixture(0, casedir)
    s = System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    for K in (1, 5, 64):
        X = np.random.default_rng(K).standard_normal((s.dim, K))
        Y = s.apply_block(torch.tensor(X, device="cuda")).cpu().numpy()
        Yf = s.apply_block(torch.tensor(X, device="cuda", dtype=torch.float32)).cpu().numpy()
        for k in range(0, K, max(1, K // 4)):
            want = o.apply(np.ascontiguousarray(X[:, k]))
            assert _rel(Y[:, k], want) < 1e-12, (K, k)
            assert _rel(Yf[:, k], want) < 2e-5, (K, k)


def test_batch_lanczos_matches_single_and_reference(casedir):
    """Lockstep batched Lanczos (graph-captured, Ritz pairs by device
    bisection / inverse iteration): FP64 alpha/beta equal the FP64 oracle's;
    FP64 energies equal the single-Hamiltonian GPU path's; ground states equal
    the reference's; every FP64 state converged to a residual below 1e-4 is
    reproduced by the mixed (FP32 vector) run within 1e-4 MeV."""
    from bigstickpublick_amd import System
    cases, d, fx, b = _batch(casedir, 8)
    r64 = b.lanczos(nkeep=5, niter=150, dtype="f64", polish=True)
    r32 = b.lanczos(nkeep=5, niter=150, dtype="f32", polish=True)
    assert r64.evals.shape == (8, 5) and r64.alpha.shape == (8, 150)
    for k in (0, 3, 6):
        c = cases[k]
        o = Oracle(*paths(c, d), c.Z, c.N, c.jz2)
        a, bb, e, _ = lanczos(o.apply, o.dim, 40, 5)
        assert np.abs(r64.alpha[k, :40] - a).max() < 1e-9, k
        assert np.abs(r64.beta[k, :40] - bb).max() < 1e-9, k
    for k, c in enumerate(cases):
        s = System(*paths(c, d), c.Z, c.N, c.jz2, upload=True)
        one = s.lanczos(nkeep=5, niter=150, dtype="f64", polish=True)
        assert np.abs(r64.evals[k] - one.evals).max() < 1e-8, k
        # rounding differences grow once Ritz values converge; early steps agree
        assert np.abs(r64.alpha[k, :60] - one.alpha[:60]).max() < 1e-10, k
        assert abs(r64.evals[k, 0] - fx["res"][k, 0, 0]) < 2e-5, k
        assert r64.resid[k, 0] < 1e-6, k
        for q in range(5):
            if r64.resid[k, q] < 1e-4:
                assert np.abs(r32.evals[k] - r64.evals[k, q]).min() < 1e-4, (k, q)


def test_batch_graph_replay(casedir):
    """The captured Lanczos graph replays to the same bits, with eager work
    and synchronisations between the replays."""
    cases, d, fx, b = _batch(casedir, 4)
    first = b.lanczos(nkeep=3, niter=60, dtype="f32", polish=True, want_vectors=True)
    for rep in range(3):
        junk = torch.randn(1 << 20, device="cuda")
        junk.mul_(2.0).sum().item()
        again = b.lanczos(nkeep=3, niter=60, dtype="f32", polish=True, want_vectors=True)
        assert np.array_equal(again.evals, first.evals), rep
        assert np.array_equal(again.alpha, first.alpha), rep
        assert torch.equal(again.evecs, first.evecs), rep


def test_batch_ritz_vectors(casedir):
    """Ritz vectors are eigenvectors of their own Hamiltonian: ||H_k x - E x||
    small, unit norm, mutually orthogonal."""
    cases, d, fx, b = _batch(casedir, 4)
    r = b.lanczos(nkeep=4, niter=150, dtype="f64", polish=True, want_vectors=True)
    X = r.evecs
    for q in range(4):
        hx = b.apply(X[q].contiguous())
        res = (hx - X[q] * torch.tensor(r.evals[:, q], device="cuda")).norm(dim=0).cpu().numpy()
        assert np.abs(res - r.resid[:, q]).max() < 1e-8
        assert res.max() < 1e-5
        assert np.abs(X[q].norm(dim=0).cpu().numpy() - 1.0).max() < 1e-12
    Xh = X.cpu().numpy()
    G = np.einsum("qik,pik->kqp", Xh, Xh)
    assert np.abs(G - np.eye(4)[None]).max() < 1e-8


def test_batch_gradient_vs_single_and_finite_differences(casedir):
    """Batched Hellmann-Feynman gradients equal the single-Hamiltonian adjoint
    pass, and central finite differences (h = 1e-4 MeV) of batched FP64
    ground-state energies: the +-h perturbed Hamiltonians are themselves one
    batch."""
    from bigstickpublick_amd import Batch, System
    cases, d, fx, b = _batch(casedir, 4)
    r = b.lanczos(nkeep=1, niter=150, dtype="f64", polish=True, want_vectors=True)
    g = b.gradient(r.evecs[0])
    assert g.shape == (4, 66)
    for k, c in enumerate(cases):
        s = System(*paths(c, d), c.Z, c.N, c.jz2, upload=True)
        gs = s.gradient(r.evecs[0][:, k].contiguous())
        assert np.abs(g[k] - gs).max() < 1e-9, k
    # finite differences for member 0: 3 SPEs and 13 TBMEs, +-h each
    base = config_case(2, 0)
    it = base.interaction
    theta = list(it.spe) + [t[6] for t in it.tbmes]
    pick = [0, 1, 2] + list(range(3, 66, 5))
    files, steps = [], []
    fd_dir = os.path.join(casedir, "cfg2_fd")
    for p in pick:
        pair = []
        for sgn in (1, -1):
            th = list(theta)
            th[p] = f32(theta[p] + sgn * 1e-4)
            c = config_case(2, 0)
            c.interaction = type(it)(it.orbits, tuple(th[:3]), [t[:6] + (v,) for t, v in zip(it.tbmes, th[3:])])
            c.name = "fd%02d%s" % (p, "p" if sgn > 0 else "m")
            write_case(c, fd_dir)
            files.append(os.path.join(fd_dir, c.name + ".int"))
            pair.append(th[p])
        steps.append(pair[0] - pair[1])
    fb = Batch(os.path.join(fd_dir, "sd.sps"), files, 4, 4, 0)
    e = fb.lanczos(nkeep=1, niter=150, dtype="f64", polish=True).evals[:, 0]
    for i, p in enumerate(pick):
        fd = (e[2 * i] - e[2 * i + 1]) / steps[i]
        assert abs(fd - g[0, p]) < 2e-5, (p, fd, g[0, p])
