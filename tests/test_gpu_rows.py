"""Spaces beyond the tile kernels' LDS: the row kernel (k_apply_rows) on
neutron sectors longer than 4096 SDs and on Delta-m operator classes larger
than 63 operators, against the oracle and the reference's own runs.

pf 1p8n (2Jz = 1, D = 219,781): neutron sectors up to 12,022 SDs, with
proton-neutron terms; pf 0p8n (D = 12,022): one neutron row of 12,022 SDs;
sdg 1p2n (2Jz = 1, D = 1,319): sd plus the 0g orbits, a Delta-m = 0 class of
110 operators.  On these only the row kernel fits; every tile variant is
refused."""
import numpy as np
import pytest
import torch

from golden_cases import paths, small_fixture
from oracle_api import Oracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.fixture(scope="module")
def wide(tmp_path_factory):
    from bigstickpublick_amd import System
    d = str(tmp_path_factory.mktemp("rows"))
    out = {}
    for name in ("pf1p8n_m1", "pf0p8n", "sdg1p2n_m1"):
        case, cd, fx = small_fixture(name, d)
        s = System(*paths(case, cd), case.Z, case.N, case.jz2, upload=True)
        o = Oracle(*paths(case, cd), case.Z, case.N, case.jz2)
        out[name] = (s, o, fx)
    yield out
    for s, _, _ in out.values():
        s.close()


@pytest.mark.parametrize("name", ["pf1p8n_m1", "pf0p8n", "sdg1p2n_m1"])
def test_row_kernel_is_the_only_variant_and_matches_oracle(name, wide):
    s, o, _ = wide[name]
    nv = len(s.variants())
    assert s.variants()[nv - 1][1] == 0  # the row kernel stages nothing
    for v in range(nv - 1):
        assert not s.set_variant("f32", v) and not s.set_variant("f64", v), v
    x = np.random.default_rng(5).standard_normal(s.dim)
    want = o.apply(x)
    assert s.set_variant("f64", nv - 1) and s.set_variant("f32", nv - 1)
    xd = torch.tensor(x, device="cuda")
    assert _rel(s.apply(xd).cpu().numpy(), want) < 1e-12
    assert _rel(s.apply(xd.float()).cpu().numpy(), want) < 2e-5
    # the autotuner finds the one variant that fits
    assert s.autotune("f32") == nv - 1 and s.autotune("f64") == nv - 1
    # y += H x
    y0 = torch.tensor(np.random.default_rng(6).standard_normal(s.dim), device="cuda")
    y = y0.clone()
    s.apply(xd, y, accumulate=True)
    assert torch.allclose(y, y0 + s.apply(xd), rtol=0, atol=1e-11)


def test_long_rows_lanczos_matches_reference_run(wide):
    """pf 1p8n: the reference's lf run (60 steps, FP32 vectors) against the
    FP64 and mixed paths: the first 20 coefficients to 2e-5 (the reference's
    FP32 rounding), the five energies to its printed digits."""
    s, _, fx = wide["pf1p8n_m1"]
    r = s.lanczos(nkeep=5, niter=60, dtype="f64", polish=True, want_vectors=True)
    print("pf1p8n FP64", r.evals, fx["res"][:, 0], r.resid)
    assert np.abs(r.alpha[:20] - fx["alpha"][:20]).max() < 2e-5
    assert np.abs(r.beta[:20] - fx["beta"][:20]).max() < 2e-5
    assert np.abs(r.evals - fx["res"][:, 0]).max() < 2e-5
    m = s.lanczos(nkeep=5, niter=60, dtype="f32", polish=True)
    assert np.abs(m.evals - fx["res"][:, 0]).max() < 1e-4
    # J of the 60-step Ritz vectors (not yet converged: the reference prints
    # J = 0.505 for the first) through the same kernel, J^2 as an operator
    for k in range(3):
        j2 = s.observable("J2", r.evecs[k])
        J = (-1.0 + np.sqrt(1.0 + 4.0 * j2)) / 2.0
        assert abs(J - fx["res"][k, 1]) < 5e-3, (k, J, fx["res"][k, 1])


def test_large_classes_spectrum(wide):
    """sdg 1p2n: FP64 Lanczos to the dense oracle spectrum and the
    reference's energies, J and T."""
    s, o, fx = wide["sdg1p2n_m1"]
    ev = np.linalg.eigvalsh(o.dense())
    r = s.lanczos(nkeep=5, niter=300, dtype="f64", polish=True, want_vectors=True)
    print("sdg1p2n FP64", r.evals, ev[:5], fx["res"][:5, 0])
    assert np.abs(r.evals - ev[:5]).max() < 1e-8
    n = min(5, len(fx["res"]))
    assert np.abs(r.evals[:n] - fx["res"][:n, 0]).max() < 2e-5
    J = fx["res"][0, 1]
    assert abs(s.observable("J2", r.evecs[0]) - J * (J + 1)) < 1e-4


def test_long_rows_block_apply(wide):
    """K vectors at once (the batched kernel without LDS) equal K single
    applies on the long-row system."""
    s, _, _ = wide["pf1p8n_m1"]
    X = torch.tensor(np.random.default_rng(7).standard_normal((s.dim, 4)), device="cuda")
    Y = s.apply_block(X)
    for k in range(4):
        y = s.apply(X[:, k].contiguous())
        assert torch.allclose(Y[:, k], y, rtol=0, atol=1e-11 * float(y.abs().max()))


@pytest.mark.parametrize("name", ["pf1p8n_m1", "sdg1p2n_m1", "pf0p8n"])
def test_gradient_beyond_lds_and_deterministic(name, wide, tmp_path):
    """The adjoint pass where its LDS staging does not fit (rows of 12,022
    SDs; sdg's NN element set): H(theta) = sum_p theta_p H_p (no mass
    scaling), so for any psi sum_p theta_p dE/dtheta_p = <psi|H|psi> /
    <psi|psi>, with <psi|H|psi> from the row kernel; and the fixed-point sums
    make a second pass bit-identical."""
    from bigstickpublick_amd.inputs import small_case
    import json, os
    from golden_cases import GOLD
    s, _, _ = wide[name]
    meta = json.load(open(os.path.join(GOLD, name + ".json")))
    case = small_case(meta["space"], meta["Z"], meta["N"], meta["jz2"], seed=meta["seed"])
    theta = list(case.interaction.spe) + [t[6] for t in case.interaction.tbmes]
    psi = torch.tensor(np.random.default_rng(8).standard_normal(s.dim), device="cuda")
    g = s.gradient(psi)
    assert len(g) == len(theta) and np.all(np.isfinite(g))
    e = float(torch.dot(psi, s.apply(psi)) / torch.dot(psi, psi))
    euler = float(np.dot(theta, g))
    assert abs(euler - e) < 1e-5 * max(1.0, abs(e)), (euler, e)
    assert np.array_equal(s.gradient(psi), g)
