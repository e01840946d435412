"""HIP path against the oracle on the GPU (through the C ABI)."""
import numpy as np
import pytest

from golden_cases import SMALL, config_fixture, paths, small_fixture
from lanczos_ref import lanczos
from oracle_api import Oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _sys(case, d):
    from bigstickpublick_amd import System
    return System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", SMALL)
def test_apply_small(name, casedir):
    case, d, fx = small_fixture(name, casedir)
    s = _sys(case, d)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(s.dim)
    want = o.apply(x)
    for v in range(len(s.variants())):
        if s.set_variant("f64", v):
            y64 = s.apply(torch.tensor(x, device="cuda")).cpu().numpy()
            assert _rel(y64, want) < 1e-12, v
        if s.set_variant("f32", v):
            y32 = s.apply(torch.tensor(x, device="cuda", dtype=torch.float32)).cpu().numpy()
            assert _rel(y32, want) < 2e-6 * 8, v


def test_accumulate_semantics(casedir):
    """y += H x (the reference's vec2 = vec2 + H vec1) versus y = H x."""
    case, d, fx = small_fixture("sd3p1n_mm2", casedir)
    s = _sys(case, d)
    rng = np.random.default_rng(2)
    x = torch.tensor(rng.standard_normal(s.dim), device="cuda")
    y0 = torch.tensor(rng.standard_normal(s.dim), device="cuda")
    y = y0.clone()
    s.apply(x, y, accumulate=True)
    hx = s.apply(x)
    assert torch.allclose(y, y0 + hx, rtol=0, atol=1e-12)


@pytest.mark.parametrize("cfg", [0, 1, 3])
def test_apply_configs(cfg, casedir):
    case, d, fx = config_fixture(cfg, casedir)
    s = _sys(case, d)
    assert s.dim == int(fx["dim"])
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    x = np.random.default_rng(3).standard_normal(s.dim)
    want = o.apply(x)
    xd = torch.tensor(x, device="cuda")
    xf = xd.float()
    for v in range(len(s.variants())):
        if s.set_variant("f64", v):
            assert _rel(s.apply(xd).cpu().numpy(), want) < 1e-12, v
        if s.set_variant("f32", v):
            assert _rel(s.apply(xf).cpu().numpy(), want) < 2e-5, v
    # the autotuner picks one of the variants and keeps it
    v = s.autotune("f32")
    assert s.get_variant("f32") == v
    assert _rel(s.apply(xf).cpu().numpy(), want) < 2e-5


def test_lanczos_fp64_config0(casedir):
    """FP64 GPU Lanczos: alpha/beta equal to the FP64 oracle's from the same
    pivot; energies equal to the reference's printed ones (f10.5)."""
    case, d, fx = config_fixture(0, casedir)
    s = _sys(case, d)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    a, b, e, _ = lanczos(o.apply, o.dim, 150, case.nkeep)
    r = s.lanczos(nkeep=case.nkeep, niter=150, dtype="f64", polish=False)
    assert np.abs(r.alpha[:60] - a[:60]).max() < 1e-9
    assert np.abs(r.beta[:60] - b[:60]).max() < 1e-9
    assert np.abs(r.evals - e).max() < 1e-8
    assert np.abs(r.evals - fx["res"][:5, 0]).max() < 1.5e-5
    # reference alpha/beta (FP32 vectors there) agree to FP32 accuracy early on
    assert np.abs(r.alpha[:20] - fx["alpha"][:20]).max() < 2e-5


def test_lanczos_mixed_vs_fp64_config0(casedir):
    """FP32 vectors + FP64 coefficients + FP64 polish: every eigenvalue within
    1e-4 MeV of the FP64 path (the stated mixed-precision target)."""
    case, d, fx = config_fixture(0, casedir)
    s = _sys(case, d)
    r64 = s.lanczos(nkeep=5, niter=150, dtype="f64", polish=True)
    r32 = s.lanczos(nkeep=5, niter=150, dtype="f32", polish=True)
    assert np.abs(r32.evals - r64.evals).max() < 1e-4
    # residuals ||Hx - Ex|| after 150 steps: the lowest states are converged
    assert np.all(r64.resid[:2] < 1e-8)
    assert np.all(r64.resid < 1e-2)


def test_thick_restart_config0(casedir):
    """Thick-restart Lanczos (keep 10, restart at 40) reaches the plain
    150-iteration FP64 energies."""
    case, d, fx = config_fixture(0, casedir)
    s = _sys(case, d)
    ref = s.lanczos(nkeep=5, niter=150, dtype="f64", polish=True)
    tr = s.lanczos(nkeep=5, niter=40, thick=10, max_restarts=30, dtype="f64", tol=1e-10, polish=True)
    # the test passes once no energy moves by 1e-10 between checks; the
    # residual is then ~1e-8 (energy error ~ residual^2 / gap)
    assert tr.resid.max() < 1e-7
    # the plain run's 4th state is only converged to a residual ~1e-3
    assert np.abs(tr.evals - ref.evals).max() < 1e-6
    assert np.abs(tr.evals[:3] - ref.evals[:3]).max() < 1e-10


def test_observables_and_host_apply(casedir):
    """J^2, T^2 of Ritz vectors are J(J+1), T(T+1) with the reference's J, T
    (.res of config 0); the host-pointer apply equals the device apply."""
    case, d, fx = config_fixture(0, casedir)
    s = _sys(case, d)
    # converged vectors; J^2 and T^2 are built, as in the reference, from FP32
    # elements (|elements| up to ~10), so <J^2> carries ~1e-6 relative error
    r = s.lanczos(nkeep=5, niter=40, thick=10, max_restarts=30, dtype="f64", tol=1e-10, polish=True,
                  want_vectors=True)
    assert r.resid[:3].max() < 1e-8
    for k in range(3):
        j2 = s.observable("J2", r.evecs[k])
        t2 = s.observable("T2", r.evecs[k])
        J, T = fx["res"][k, 1], fx["res"][k, 2]
        assert abs(j2 - J * (J + 1)) < 5e-5, (k, j2, J)
        assert abs(t2 - T * (T + 1)) < 5e-5, (k, t2, T)
    x = np.random.default_rng(9).standard_normal(s.dim)
    yd = s.apply(torch.tensor(x, device="cuda")).cpu().numpy()
    yh = s.apply_host(x)
    assert np.array_equal(yd, yh)


def test_operator_matrix(casedir):
    """<v_f|O|v_i> of an operator given as an .int file (the reference's
    expectation-value path): equal to the oracle's apply of the same operator;
    with O = H on converged Ritz vectors, the diagonal matrix of the energies."""
    import os
    from bigstickpublick_amd.inputs import small_case
    case, d, fx = small_fixture("sd2p2n", casedir)
    s = _sys(case, d)
    op = small_case("sd", 2, 2, 0, seed=5).interaction
    opf = os.path.join(d, "op5.int")
    with open(opf, "w") as f:
        f.write(op.int_text())
    rng = np.random.default_rng(12)
    V = rng.standard_normal((3, s.dim))
    M = s.operator_matrix(opf, torch.tensor(V, device="cuda"))
    o = Oracle(paths(case, d)[0], opf, case.Z, case.N, case.jz2)
    want = np.array([[V[f] @ o.apply(V[i]) for f in range(3)] for i in range(3)])
    assert _rel(M, want) < 1e-12
    M32 = s.operator_matrix(opf, torch.tensor(V, device="cuda", dtype=torch.float32))
    assert _rel(M32, want) < 1e-5
    r = s.lanczos(nkeep=3, niter=min(s.dim, 300), dtype="f64", polish=True, want_vectors=True)
    H = s.operator_matrix(paths(case, d)[1], r.evecs)
    assert np.abs(np.diag(H) - r.evals).max() < 1e-8
    assert np.abs(H - np.diag(np.diag(H))).max() < 1e-8


@pytest.mark.gpu
def test_states_from_reference_wfn(casedir):
    """The reference's sd2p2n.wfn read onto the device through a system built
    from the file: its own Hamiltonian (intf None) on those states is
    diagonal with the stored energies (FP32 vectors, converged to 1e-4), and
    a second operator's matrix equals the oracle's on the same vectors."""
    import os
    from bigstickpublick_amd.engine import System
    from bigstickpublick_amd.inputs import small_case
    case, d, fx = small_fixture("sd2p2n", casedir)
    wfn = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sd2p2n.wfn")
    w = System.from_wfn(wfn, paths(case, d)[1], upload=True)
    e, j, t, v = w.read_wfn(wfn, dtype="f64", device=True)
    H = w.operator_matrix(None, v)
    vh = v.cpu().numpy()
    nrm = np.einsum("ij,ij->i", vh, vh)
    assert np.abs(np.diag(H) / nrm - e).max() < 1e-4
    assert np.abs(H - np.diag(np.diag(H))).max() < 1e-4
    op = small_case("sd", 2, 2, 0, seed=5).interaction
    opf = os.path.join(d, "op5.int")
    with open(opf, "w") as f:
        f.write(op.int_text())
    M = w.operator_matrix(opf, v)
    o = Oracle(paths(case, d)[0], opf, case.Z, case.N, case.jz2)
    want = np.array([[vh[f] @ o.apply(vh[i]) for f in range(len(e))] for i in range(len(e))])
    assert _rel(M, want) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,tol", [("f64", 1e-10), ("f32", 2e-5)])
def test_strength_function_moments(dtype, tol, casedir):
    """Lanczos from a pivot p is a Gauss quadrature of p's spectral measure:
    sum_k S_k E_k^n = <p|H^n|p> for n < 2 niter.  Checked for n = 0..8
    against the oracle's H (tolerance relative to |<p|H^n|p>| + ||p||^2 ||H||^n),
    and the pivot norm is ||p||."""
    import torch
    case, d, fx = small_fixture("sd2p2n", casedir)
    s = _sys(case, d)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    p = np.random.default_rng(21).standard_normal(s.dim) * 0.05
    tdt = torch.float64 if dtype == "f64" else torch.float32
    r = s.strength(torch.tensor(p, dtype=tdt, device="cuda"), nkeep=4, niter=24)
    pq = p.astype(np.float32).astype(np.float64) if dtype == "f32" else p
    assert abs(r.pivot_norm - np.linalg.norm(pq)) <= 1e-12 * np.linalg.norm(pq)
    assert (r.strength >= 0).all() and np.all(np.diff(r.ritz) > 0)
    hn = max(abs(r.ritz[0]), abs(r.ritz[-1]))
    v = pq.copy()
    for n in range(9):
        want = float(pq @ v)
        got = float(np.sum(r.strength * r.ritz ** n))
        assert abs(got - want) <= tol * (abs(want) + pq @ pq * hn ** n), (n, got, want)
        v = o.apply(v)
    # the lowest Ritz values equal bs_lanczos's from the same pivot direction
    assert np.abs(r.evals[:2] - r.ritz[:2]).max() < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("block", [1, 3, 4, 8])
def test_block_lanczos_small(block, casedir):
    """Block Lanczos on sd 2p2n (D = 640, a basis of 240 or 400 vectors,
    past the point where the lowest states converge): the lowest five equal
    the oracle's dense spectrum (FP64 1e-9, FP32 vectors with FP64 polish
    1e-5), with FP64 residuals below 1e-6, and the Ritz vectors are
    orthonormal."""
    import torch
    case, d, fx = small_fixture("sd2p2n", casedir)
    s = _sys(case, d)
    ev = np.linalg.eigvalsh(Oracle(*paths(case, d), case.Z, case.N, case.jz2).dense())[:5]
    n = 240 if block <= 4 else 400
    r = s.block_lanczos(nkeep=5, niter=n, block=block, dtype="f64", want_vectors=True)
    assert np.abs(r.evals - ev).max() < 1e-9, (r.evals, ev)
    assert r.resid.max() < 1e-6
    X = r.evecs.cpu().numpy()
    assert np.abs(X @ X.T - np.eye(5)).max() < 1e-10
    r32 = s.block_lanczos(nkeep=5, niter=n, block=block, dtype="f32")
    assert np.abs(r32.evals - ev).max() < 1e-5


@pytest.mark.gpu
def test_block_strength_moments(casedir):
    """bs_block_strength on sd 2p2n from three seeded pivots, 6 blocks: the
    strengths of each pivot sum to |p|^2 and reproduce its moments
    sum_k S_ik E_k^n = <p_i|H^n|p_i> for n <= 2 * nblocks - 1 (Gauss block
    quadrature) against the oracle's dense H, in FP64; the FP32 run gives
    the same Ritz values to 1e-4."""
    import torch
    case, d, fx = small_fixture("sd2p2n", casedir)
    s = _sys(case, d)
    H = Oracle(*paths(case, d), case.Z, case.N, case.jz2).dense()
    rng = np.random.default_rng(11)
    P = rng.standard_normal((3, s.dim))
    ev, st, pn2, mv = s.block_strength(torch.tensor(P, device="cuda"), 6, dtype="f64")
    assert mv == 18
    assert np.abs(pn2 - (P * P).sum(1)).max() < 1e-9 * pn2.max()
    assert np.abs(st.sum(1) - pn2).max() < 1e-9 * pn2.max()
    for i in range(3):
        v = P[i].copy()
        for n in range(12):
            mom = P[i] @ v
            scale = (st[i] * np.abs(ev) ** n).sum()
            assert abs((st[i] * ev ** n).sum() - mom) < 1e-9 * max(1.0, scale), (i, n)
            v = H @ v
    ev32, st32, _, _ = s.block_strength(torch.tensor(P, device="cuda", dtype=torch.float32), 6, dtype="f32")
    assert np.abs(ev32 - ev).max() < 1e-4


@pytest.mark.gpu
def test_block_lanczos_config0(casedir):
    """sd 4p4n (D = 28,503): block 4 with 300 basis vectors gives the lowest
    five of a 300-step single-vector FP64 run; with tol it stops earlier at
    the same energies."""
    case, d, fx = config_fixture(0, casedir)
    s = _sys(case, d)
    single = s.lanczos(nkeep=5, niter=300, dtype="f64")
    blk = s.block_lanczos(nkeep=5, niter=300, block=4, dtype="f64")
    assert np.abs(blk.evals - single.evals).max() < 1e-7, (blk.evals, single.evals)
    early = s.block_lanczos(nkeep=5, niter=300, block=4, dtype="f64", tol=1e-9)
    assert s.block_lanczos(nkeep=5, niter=300, block=4, dtype="f32").evals.shape == (5,)
    assert early.matvecs < blk.matvecs and np.abs(early.evals - single.evals).max() < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sd2p2n", "pf2p1n_m3", "sd1p0n_m1"])
def test_exact_diagonalization(name, casedir):
    """Option ex: the dense H from 64-column blocks and rocSOLVER give the
    oracle's dense spectrum (all states, 1e-10) with eigenvector residuals at
    rounding level, FP32 output vectors equal to the FP64 ones rounded."""
    case, d, fx = small_fixture(name, casedir)
    s = _sys(case, d)
    ev = np.linalg.eigvalsh(Oracle(*paths(case, d), case.Z, case.N, case.jz2).dense())
    n = min(10, s.dim)
    r = s.exact(nkeep=n, want_vectors=True)
    assert r.matvecs == s.dim
    assert np.abs(r.evals - ev[:n]).max() < 1e-10, (r.evals, ev[:n])
    assert r.resid.max() < 1e-10
    X = r.evecs.cpu().numpy()
    assert np.abs(X @ X.T - np.eye(n)).max() < 1e-12
    r32 = s.exact(nkeep=n, dtype="f32", want_vectors=True)
    assert np.abs(r32.evals - r.evals).max() == 0.0
    x32 = r32.evecs.cpu().numpy().astype(np.float64)
    ok = np.abs(np.abs(np.einsum("ij,ij->i", x32, X)) - 1.0)
    assert ok[np.diff(np.r_[r.evals, np.inf]) > 1e-6].max() < 1e-6


@pytest.mark.gpu
def test_plain_lanczos_stops_at_tol(casedir):
    """A plain run with a tolerance (options ld / lc) checks its Ritz values
    every 10 steps and stops once none moves by tol: fewer matvecs than the
    iteration cap, the same energies as the full run."""
    case, d, fx = config_fixture(0, casedir)
    s = _sys(case, d)
    full = s.lanczos(nkeep=5, niter=300, dtype="f64")
    early = s.lanczos(nkeep=5, niter=300, dtype="f64", tol=1e-8)
    assert early.matvecs < full.matvecs and early.matvecs % 10 == 5  # 10k steps + 5 polish applies
    assert np.abs(early.evals - full.evals).max() < 1e-6
    assert np.count_nonzero(early.alpha) == early.matvecs - 5
