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
    want = [Oracle(*paths(c, d), c.Z, c.N, c.jz2).apply(np.ascontiguousarray(X[:, k]))
            for k, c in enumerate(cases)]
    xd = torch.tensor(X, device="cuda")
    for v in range(5):
        assert b.set_variant("f64", v) and b.set_variant("f32", v), v
        Y64 = b.apply(xd).cpu().numpy()
        Y32 = b.apply(xd.float()).cpu().numpy()
        for k in range(b.K):
            assert _rel(Y64[:, k], want[k]) < 1e-12, (v, k)
            assert _rel(Y32[:, k], want[k]) < 2e-5, (v, k)


def test_batch_two_row_tiles_match_oracle(casedir):
    """The two-row K-vector tiles (variants 5, 6) with one Hamiltonian per
    vector: four members of the sd 4p4n ensemble, FP32 at K = 4 and FP64 at
    K = 2, against the oracle."""
    from bigstickpublick_amd import Batch
    cases, d, fx = ensemble_fixture(casedir, 4)
    for K, dt, tol in ((4, "f32", 2e-5), (2, "f64", 1e-12)):
        b = Batch(os.path.join(d, "sd.sps"), [paths(c, d)[1] for c in cases[:K]], 4, 4, 0)
        X = np.random.default_rng(40 + K).standard_normal((b.dim, K))
        want = [Oracle(*paths(c, d), c.Z, c.N, c.jz2).apply(np.ascontiguousarray(X[:, k]))
                for k, c in enumerate(cases[:K])]
        xd = torch.tensor(X, device="cuda", dtype=torch.float32 if dt == "f32" else torch.float64)
        for v in (5, 6):
            assert b.set_variant(dt, v), (K, v)
            Y = b.apply(xd).cpu().numpy()
            for k in range(K):
                assert _rel(Y[:, k], want[k]) < tol, (K, v, k)
        b.close()


def test_batch_apply_pf_matches_oracle(casedir):
    """Two pf Hamiltonians on the pf 4p4n basis (D = 1,963,461), every
    multi-vector variant, against the oracle."""
    from bigstickpublick_amd import Batch
    from bigstickpublick_amd.inputs import PF_ORBITS, PF_SPE, gaussian_interaction
    case, d, fx = config_fixture(3, casedir)
    other = config_case(3)
    other.interaction = gaussian_interaction(PF_ORBITS, PF_SPE, seed=3001)
    other.name = "pf4p4n_b"
    write_case(other, d)
    cases = [case, other]
    b = Batch(os.path.join(d, "pf.sps"), [paths(c, d)[1] for c in cases], 4, 4, 0)
    X = np.random.default_rng(9).standard_normal((b.dim, 2))
    want = [Oracle(*paths(c, d), c.Z, c.N, c.jz2).apply(np.ascontiguousarray(X[:, k]))
            for k, c in enumerate(cases)]
    xd = torch.tensor(X, device="cuda")
    for v in range(7):
        if b.set_variant("f64", v):
            Y64 = b.apply(xd).cpu().numpy()
            for k in range(2):
                assert _rel(Y64[:, k], want[k]) < 1e-12, (v, k)
        assert b.set_variant("f32", v), v
        Y32 = b.apply(xd.float()).cpu().numpy()
        for k in range(2):
            assert _rel(Y32[:, k], want[k]) < 2e-5, (v, k)


def test_block_apply_matches_single(casedir):
    """Same H on K interleaved vectors (the reference's block apply) equals the
    oracle's single applies, for every multi-vector variant."""
    from bigstickpublick_amd import System
    for cfg in (0, 3):
        case, d, fx = config_fixture(cfg, casedir)
        s = System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)
        o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
        for K in ((1, 2, 4, 8, 5, 64) if cfg == 0 else (1, 4)):
            X = np.random.default_rng(K).standard_normal((s.dim, K))
            cols = list(range(0, K, max(1, K // 4)))
            want = {k: o.apply(np.ascontiguousarray(X[:, k])) for k in cols}
            xd = torch.tensor(X, device="cuda")
            for v in range(7):
                if s.set_block_variant("f64", K, v):
                    Y = s.apply_block(xd).cpu().numpy()
                    for k in cols:
                        assert _rel(Y[:, k], want[k]) < 1e-12, (cfg, K, v, k)
                if s.set_block_variant("f32", K, v):
                    Yf = s.apply_block(xd.float()).cpu().numpy()
                    for k in cols:
                        assert _rel(Yf[:, k], want[k]) < 2e-5, (cfg, K, v, k)


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
        a, bb, e, _ = lanczos(o.apply, o.dim, 30, 5)
        assert np.abs(r64.alpha[k, :30] - a).max() < 1e-10, k
        assert np.abs(r64.beta[k, :30] - bb).max() < 1e-10, k
    for k, c in enumerate(cases):
        s = System(*paths(c, d), c.Z, c.N, c.jz2, upload=True)
        one = s.lanczos(nkeep=5, niter=150, dtype="f64", polish=True)
        assert np.abs(r64.evals[k] - one.evals).max() < 1e-8, k
        # rounding differences grow ~3x per step once a Ritz value has
        # converged (from step ~45 here); the early steps agree to rounding
        assert np.abs(r64.alpha[k, :30] - one.alpha[:30]).max() < 1e-10, k
        assert abs(r64.evals[k, 0] - fx["res"][k, 0, 0]) < 2e-5, k
        assert r64.resid[k, 0] < 1e-6, k
        for q in range(5):
            if r64.resid[k, q] < 1e-4:
                assert np.abs(r32.evals[k] - r64.evals[k, q]).min() < 1e-4, (k, q)


def test_batch_thick_restart_matches_single(casedir):
    """Thick restart (keep 10, restart at 40) for all Hamiltonians in
    lockstep, arrowhead Ritz solve on the device, against the single-
    Hamiltonian path (host Jacobi) run to convergence, FP64 and mixed."""
    from bigstickpublick_amd import System
    cases, d, fx, b = _batch(casedir, 4)
    r = b.lanczos(nkeep=5, niter=40, thick=10, max_restarts=60, dtype="f64", tol=1e-10, polish=True)
    rm = b.lanczos(nkeep=5, niter=40, thick=10, max_restarts=60, dtype="f32", tol=1e-6, polish=True)
    assert r.matvecs > 40
    for k, c in enumerate(cases):
        s = System(*paths(c, d), c.Z, c.N, c.jz2, upload=True)
        one = s.lanczos(nkeep=5, niter=40, thick=10, max_restarts=60, dtype="f64", tol=1e-10, polish=True)
        conv = (one.resid < 1e-5) & (r.resid[k] < 1e-5)
        assert conv[0], (k, one.resid, r.resid[k])
        assert np.abs(r.evals[k][conv] - one.evals[conv]).max() < 1e-8, k
        assert np.abs(rm.evals[k][conv] - one.evals[conv]).max() < 1e-4, k
        assert abs(r.evals[k, 0] - fx["res"][k, 0, 0]) < 2e-5, k


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
    """Ritz vectors: the reported residuals are ||H_k x - E x|| of the
    returned vectors, ground states converge, vectors are orthonormal."""
    cases, d, fx, b = _batch(casedir, 4)
    r = b.lanczos(nkeep=4, niter=150, dtype="f64", polish=True, want_vectors=True)
    X = r.evecs
    for q in range(4):
        hx = b.apply(X[q].contiguous())
        res = (hx - X[q] * torch.tensor(r.evals[:, q], device="cuda")).norm(dim=0).cpu().numpy()
        assert np.abs(res - r.resid[:, q]).max() < 1e-8
        if q == 0:  # excited states need not converge in 150 steps
            assert res.max() < 1e-6
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
    # finite differences for member 0: 3 SPEs and 13 TBMEs, +-h each, at
    # h = 1e-4 MeV and h = 1e-2 MeV, all 64 perturbed Hamiltonians in one
    # batch.  The matrix elements are FP32 (as in the reference), so E(theta)
    # carries rounding steps of ~1e-8 MeV whatever the vector precision: at
    # h = 1e-4 they make the quotient uncertain by up to ~1e-3 (the oracle
    # gives the same quotients), at h = 1e-2 by ~1e-6, where the truncation
    # error h^2 E^(3) / 6 is still below 1e-5.
    base = config_case(2, 0)
    it = base.interaction
    theta = list(it.spe) + [t[6] for t in it.tbmes]
    pick = [0, 1, 2] + list(range(3, 66, 5))
    files, steps = [], []
    fd_dir = os.path.join(casedir, "cfg2_fd")
    for h in (1e-4, 1e-2):
        for p in pick:
            pair = []
            for sgn in (1, -1):
                th = list(theta)
                th[p] = f32(theta[p] + sgn * h)
                c = config_case(2, 0)
                c.interaction = type(it)(it.orbits, tuple(th[:3]),
                                         [t[:6] + (v,) for t, v in zip(it.tbmes, th[3:])])
                c.name = "fd%02d%s%d" % (p, "p" if sgn > 0 else "m", int(h > 1e-3))
                write_case(c, fd_dir)
                files.append(os.path.join(fd_dir, c.name + ".int"))
                pair.append(th[p])
            steps.append(pair[0] - pair[1])
    assert len(files) == 64
    fb = Batch(os.path.join(fd_dir, "sd.sps"), files, 4, 4, 0)
    e = fb.lanczos(nkeep=1, niter=150, dtype="f64", polish=True).evals[:, 0]
    n = len(pick)
    for i, p in enumerate(pick):
        fd4 = (e[2 * i] - e[2 * i + 1]) / steps[i]
        fd2 = (e[2 * (n + i)] - e[2 * (n + i) + 1]) / steps[n + i]
        assert abs(fd4 - g[0, p]) < 2e-3, (p, fd4, g[0, p])
        assert abs(fd2 - g[0, p]) < 5e-5, (p, fd2, g[0, p])


def test_batch_set_interactions_equals_fresh_batch(casedir):
    """New values on the kept basis and jump tables: members 4..7 of the
    ensemble put into a batch built from members 0..3 give the same H x,
    bit for bit, and the same Lanczos energies as a batch built from 4..7;
    a file that cannot be read leaves the previous values in place."""
    from bigstickpublick_amd import Batch
    cases, d, fx = ensemble_fixture(casedir, 8)
    first = [paths(c, d)[1] for c in cases[:4]]
    second = [paths(c, d)[1] for c in cases[4:]]
    b = Batch(os.path.join(d, "sd.sps"), first, 4, 4, 0)
    fresh = Batch(os.path.join(d, "sd.sps"), second, 4, 4, 0)
    x = torch.tensor(np.random.default_rng(12).standard_normal((b.dim, 4)), device="cuda", dtype=torch.float32)
    y_first = b.apply(x)
    r_first = b.lanczos(nkeep=2, niter=120, dtype="f64", polish=True)
    b.set_interactions(second)
    assert torch.equal(b.apply(x), fresh.apply(x))
    r = b.lanczos(nkeep=2, niter=120, dtype="f64", polish=True)
    want = fresh.lanczos(nkeep=2, niter=120, dtype="f64", polish=True)
    print("set_interactions", r.evals, want.evals)
    assert np.abs(r.evals - want.evals).max() < 1e-10
    assert np.abs(r.evals - r_first.evals).max() > 1e-3  # different Hamiltonians
    with pytest.raises(RuntimeError):
        b.set_interactions(first[:3] + [os.path.join(d, "missing.int")])
    assert torch.equal(b.apply(x), fresh.apply(x))
    b.set_interactions(first)
    assert torch.equal(b.apply(x), y_first)
    with pytest.raises(ValueError):
        b.set_interactions(first[:3])
    b.close()
    fresh.close()
