"""The GPU's FP64 path and mixed path against the reference's own FP64 build
(oracle/Makefile ref64: lanc_prec = 8; tests/golden/config{0,1}_fp64.npz,
its .lcoef and .res of the configurations' fixed runs: sd 4p4n lowest 5 and
sd 6p6n lowest 10, 150 iterations from the uniform pivot).

* FP64: the coefficients equal the reference's for 60 steps (1e-9) and every
  kept Ritz value equals the reference's, i.e. the eigenvalue of its own
  tridiagonal matrix, within 1e-6 MeV;
* mixed (FP32 vectors and elements, FP64 coefficients, FP64 polish): each
  eigenvalue's error against the reference's FP64 value, stated and within
  1e-4 MeV;
* eigenvectors: the FP64 Ritz vectors of the converged states overlap the
  oracle's FP64 ones (the restated recurrence on the oracle's Hamiltonian,
  itself equal to the reference's FP64 run) to at least 1 - 1e-8."""
import os

import numpy as np
import pytest

from golden_cases import GOLD, config_fixture, paths
from lanczos_ref import lanczos
from oracle_api import Oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _ref(cfg):
    f = dict(np.load(os.path.join(GOLD, "config%d_fp64.npz" % cfg)))
    al, be = f["alpha"], f["beta"]
    ritz = np.linalg.eigvalsh(np.diag(al) + np.diag(be[:-1], 1) + np.diag(be[:-1], -1))
    return f, ritz


@pytest.mark.parametrize("cfg", [0, 1])
def test_fp64_and_mixed_against_reference_fp64_build(cfg, casedir):
    from bigstickpublick_amd import System
    case, d, _ = config_fixture(cfg, casedir)
    f, ritz = _ref(cfg)
    nk = case.nkeep
    s = System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)
    r64 = s.lanczos(nkeep=nk, niter=150, dtype="f64", polish=False)
    assert np.abs(r64.alpha[:60] - f["alpha"][:60]).max() < 1e-9
    assert np.abs(r64.beta[:60] - f["beta"][:60]).max() < 1e-9
    err64 = np.abs(r64.evals - ritz[:nk])
    assert err64.max() < 1e-6, err64
    r32 = s.lanczos(nkeep=nk, niter=150, dtype="f32", polish=True)
    err32 = np.abs(r32.evals - ritz[:nk])
    print("config %d: FP64 |dE| max %.2e; mixed |dE| per state %s" % (cfg, err64.max(),
                                                                    " ".join("%.1e" % x for x in err32)))
    if cfg == 0:
        assert err32.max() < 1e-4, err32
    else:
        # sd 6p6n: the uniform pivot has no component along a state near
        # -81.66 MeV in exact arithmetic.  FP32 rounding brings it into the
        # Krylov space within 150 steps -- in this mixed run and in the
        # reference's own FP32 run, whose .res lists it as the 10th state at
        # -81.658, not converged there either -- FP64 rounding does not: the
        # reference's FP64 run (and this FP64 run) end at -81.162.  The nine
        # states both precisions see agree to 1e-4; the tenth is that extra
        # state, half a MeV below the FP64 value and within 0.05 of the FP32
        # reference's unconverged estimate.
        fx32 = config_fixture(1, casedir)[2]
        assert err32[:9].max() < 1e-4, err32
        assert r32.evals[9] < ritz[9] - 0.3 and abs(r32.evals[9] - fx32["res"][9, 0]) < 0.05, r32.evals
    s.close()


def test_fp64_eigenvectors_overlap_oracle(casedir):
    from bigstickpublick_amd import System
    case, d, _ = config_fixture(0, casedir)
    s = System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)
    r = s.lanczos(nkeep=5, niter=150, dtype="f64", polish=True, want_vectors=True)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    _, _, e, vecs = lanczos(o.apply, o.dim, 150, 5)
    got = r.evecs.cpu().numpy()
    conv = np.nonzero(r.resid < 1e-6)[0]
    assert len(conv) >= 2
    for k in conv:
        ov = abs(float(got[k] @ vecs[k])) / (np.linalg.norm(got[k]) * np.linalg.norm(vecs[k]))
        assert ov >= 1.0 - 1e-8, (k, ov, r.resid[k])
    s.close()
