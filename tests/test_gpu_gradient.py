"""Hellmann-Feynman gradients from the adjoint pass against central finite
differences (h = 1e-4 MeV) of FP64 eigenvalues."""
import os

import numpy as np
import pytest

from bigstickpublick_amd.inputs import f32, small_case, write_case
from golden_cases import config_fixture, paths
from oracle_api import Oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _energy_identity(s, psi):
    """sum_me v_me rho_me must equal <psi|H|psi> (all three operator classes)."""
    rpp, rnn, rpn, n2 = s.density(psi)
    e = ((s.me_table(0) * rpp).sum() + (s.me_table(1) * rnn).sum() + (s.me_table(2) * rpn).sum()) / n2
    hx = s.apply(psi)
    return e, float((psi.double() @ hx.double()).item() / n2)


def test_density_reproduces_energy(casedir):
    case, d, fx = config_fixture(0, casedir)
    from bigstickpublick_amd import System
    s = System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)
    psi = torch.tensor(np.random.default_rng(4).standard_normal(s.dim), device="cuda")
    e1, e2 = _energy_identity(s, psi)
    assert abs(e1 - e2) < 1e-9 * max(1.0, abs(e2))
    e1, e2 = _energy_identity(s, psi.float())
    assert abs(e1 - e2) < 1e-4 * max(1.0, abs(e2))


def _oracle_h(case, d, theta, tag):
    """Dense oracle Hamiltonian of the interaction with parameters theta."""
    norb = len(case.orbits)
    it = case.interaction
    c = type(case)(**{**case.__dict__})
    c.interaction = type(it)(it.orbits, tuple(theta[:norb]), [l[:6] + (v,) for l, v in zip(it.tbmes, theta[norb:])])
    c.name = tag
    sps, intf = write_case(c, d)
    return Oracle(os.path.join(d, sps + ".sps"), os.path.join(d, intf + ".int"), case.Z, case.N, case.jz2).dense()


@pytest.mark.parametrize("space,Z,N,jz2", [("sd", 2, 2, 0), ("sd", 1, 2, 1), ("pf", 2, 1, 3), ("sd", 1, 0, 1)])
def test_gradient_vs_finite_differences(space, Z, N, jz2, casedir):
    """H is linear in the parameters (SPEs, TBMEs) up to the FP32 rounding of
    its elements, so dE0/dtheta_p = <psi|H(e_p)|psi> with H(e_p) the oracle's
    Hamiltonian of the unit interaction e_p (Hellmann-Feynman, exact for the
    linearised H): the adjoint pass must reproduce it to rounding.  Central
    differences of the oracle's E0 check the whole chain at h = 1e-2 MeV;
    at h = 1e-4 MeV the FP32 elements (steps of ~5e-7 in elements of a few
    MeV) make the quotient uncertain by up to ~1e-2, so h = 1e-4 is reported
    against a bound of that size only."""
    from bigstickpublick_amd import System
    case = small_case(space, Z, N, jz2, seed=91)
    d = os.path.join(casedir, "grad_%s%d%d" % (space, Z, N))
    write_case(case, d)
    s = System(*paths(case, d), Z, N, jz2, upload=True)
    r = s.lanczos(nkeep=1, niter=min(s.dim, 200), dtype="f64", polish=True, want_vectors=True)
    g = s.gradient(r.evecs[0])
    theta = list(case.interaction.spe) + [t[6] for t in case.interaction.tbmes]
    nspe, ntb = s.num_params()
    assert len(g) == nspe + ntb == len(theta)
    H0 = _oracle_h(case, d, theta, "h0")
    w, V = np.linalg.eigh(H0)
    psi = V[:, 0]
    assert abs(w[0] - r.evals[0]) < 1e-9
    rng = np.random.default_rng(0)
    pick = list(range(nspe)) + sorted(rng.choice(np.arange(nspe, len(theta)), size=min(12, ntb), replace=False))
    for p in pick:
        unit = [0.0] * len(theta)
        unit[p] = 1.0
        hf = float(psi @ _oracle_h(case, d, unit, "u") @ psi)
        assert abs(hf - g[p]) < 1e-8, (p, hf, g[p])
        for h, tol in ((1e-2, 1e-4), (1e-4, 2e-2)):
            tp, tm = list(theta), list(theta)
            tp[p] = f32(theta[p] + h)
            tm[p] = f32(theta[p] - h)
            ep = np.linalg.eigvalsh(_oracle_h(case, d, tp, "p"))[0]
            em = np.linalg.eigvalsh(_oracle_h(case, d, tm, "m"))[0]
            fd = (ep - em) / (tp[p] - tm[p])
            assert abs(fd - g[p]) < tol, (p, h, fd, g[p])


def test_autograd_function_config0(casedir):
    """torch.autograd through EigenEnergies: the gradient of E0 + E1 equals the
    sum of the adjoint gradients of the two eigenvectors (computed here
    independently), and central differences of the GPU FP64 energies at
    h = 1e-2 MeV.  With FP32 matrix elements E(theta) is a step function at
    the 1e-7 level, which puts ~4e-2 / (h / 1e-4) of rounding noise into the
    quotient for this system: ~4e-4 at h = 1e-2, hence the bound."""
    from bigstickpublick_amd import System
    from bigstickpublick_amd.autograd import eigenenergies, parameters
    case, d, fx = config_fixture(0, casedir)
    theta = parameters(case.interaction).requires_grad_(True)
    e = eigenenergies(theta, case.interaction, case.Z, case.N, case.jz2, nkeep=2, niter=150, dtype="f64")
    assert np.abs(e.detach().numpy() - fx["res"][:2, 0]).max() < 2e-5
    (e[0] + e[1]).backward()
    g = theta.grad.numpy()
    s = System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)
    r = s.lanczos(nkeep=2, niter=150, dtype="f64", polish=True, want_vectors=True)
    want = s.gradient(r.evecs[0]) + s.gradient(r.evecs[1])
    # the batched solve's Ritz vectors (bisection + inverse iteration) against
    # the single solve's (Jacobi): equal to FP64 rounding
    print("autograd vs single adjoint", np.abs(g - want).max())
    assert np.abs(g - want).max() < 1e-8
    for p in (0, 2, 5, 40):
        tp = theta.detach().clone()
        tm = theta.detach().clone()
        tp[p] = f32(float(tp[p]) + 1e-2)
        tm[p] = f32(float(tm[p]) - 1e-2)
        ep = eigenenergies(tp, case.interaction, case.Z, case.N, case.jz2, nkeep=2, niter=150, dtype="f64")
        em = eigenenergies(tm, case.interaction, case.Z, case.N, case.jz2, nkeep=2, niter=150, dtype="f64")
        fd = float((ep.sum() - em.sum()) / (tp[p] - tm[p]))
        assert abs(fd - g[p]) < 1.5e-3, (p, fd, g[p])


def test_ensemble_autograd_reuses_tables(casedir):
    """ensemble_energies over three members of the sd 4p4n ensemble: the
    energies equal each member's own FP64 solve, the gradient of a weighted
    sum equals the weighted single-Hamiltonian adjoint gradients, and a
    second call with other parameters reuses the cached basis and jump tables
    (same batch object) while giving the new Hamiltonians' energies."""
    from bigstickpublick_amd import System, autograd
    from bigstickpublick_amd.autograd import ensemble_energies, parameters
    from golden_cases import ensemble_fixture
    autograd.clear_cache()
    cases, d, fx = ensemble_fixture(casedir, 6)
    tmpl = cases[0].interaction
    thetas = torch.stack([parameters(c.interaction) for c in cases[:3]]).requires_grad_(True)
    e = ensemble_energies(thetas, tmpl, 4, 4, 0, nkeep=2, niter=150, dtype="f64")
    assert e.shape == (3, 2)
    w = torch.tensor([[1.0, 0.5], [-2.0, 0.0], [0.25, 1.5]], dtype=torch.float64)
    (e * w).sum().backward()
    g = thetas.grad.numpy()
    for k, c in enumerate(cases[:3]):
        s = System(*paths(c, d), c.Z, c.N, c.jz2, upload=True)
        r = s.lanczos(nkeep=2, niter=150, dtype="f64", polish=True, want_vectors=True)
        assert np.abs(e[k].detach().numpy() - r.evals).max() < 1e-8, (k, e[k], r.evals)
        want = float(w[k, 0]) * s.gradient(r.evecs[0]) + float(w[k, 1]) * s.gradient(r.evecs[1])
        print("member", k, np.abs(g[k] - want).max())
        assert np.abs(g[k] - want).max() < 1e-8, k
        s.close()
    assert len(autograd._cache) == 1
    b = next(iter(autograd._cache.values()))
    other = torch.stack([parameters(c.interaction) for c in cases[3:6]])
    e2 = ensemble_energies(other, tmpl, 4, 4, 0, nkeep=2, niter=150, dtype="f64")
    assert len(autograd._cache) == 1 and next(iter(autograd._cache.values())) is b
    for k, c in enumerate(cases[3:6]):
        s = System(*paths(c, d), c.Z, c.N, c.jz2, upload=True)
        r = s.lanczos(nkeep=2, niter=150, dtype="f64", polish=True)
        assert np.abs(e2[k].numpy() - r.evals).max() < 1e-8, (k, e2[k], r.evals)
        s.close()
    autograd.clear_cache()
