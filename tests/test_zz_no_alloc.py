"""bs_lanczos_noalloc (the Lanczos entry point that allocates no device
memory: its vectors live in a fixed device scratch array) against bs_lanczos,
through the C ABI: two calls in a row on the same GPU, the second on a
smaller basis, each equal to the allocating entry point's run; a request
larger than the scratch is refused with an error that states the sizes."""
import ctypes

import numpy as np
import pytest

from golden_cases import config_fixture, paths

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _noalloc(s, nkeep, niter, dtype, tol=0.0):
    from bigstickpublick_amd.engine import BS_F32, BS_F64, _Params, load_library
    L = load_library()
    p = _Params(nkeep, niter, 0, 0, BS_F32 if dtype == "f32" else BS_F64, tol, 0)
    m = min(niter, s.dim)
    a, b, e = np.zeros(m), np.zeros(m), np.zeros(nkeep)
    it = ctypes.c_int()
    rc = L.bs_lanczos_noalloc(s.h, ctypes.byref(p), a.ctypes.data, b.ctypes.data, e.ctypes.data, ctypes.byref(it))
    return rc, a, b, e, it.value, L.bs_last_error().decode()


def test_noalloc_matches_allocating_entry_point(casedir):
    from bigstickpublick_amd import System
    runs = []
    for cfg, dtype, nkeep, niter in ((1, "f32", 10, 150), (0, "f64", 5, 120), (0, "f32", 5, 60)):
        case, d, fx = config_fixture(cfg, casedir)
        s = System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)
        rc, a, b, e, it, msg = _noalloc(s, nkeep, niter, dtype)
        assert rc == 0, msg
        ref = s.lanczos(nkeep=nkeep, niter=niter, dtype=dtype, polish=False)
        # (the allocating run also applies H once per kept vector for the
        # residuals the binding asks for)
        assert it == niter and ref.matvecs == niter + nkeep
        for got, want in ((a, ref.alpha), (b, ref.beta), (e, ref.evals)):
            assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), (cfg, dtype)
        runs.append((cfg, e))
        s.close()
    # the reference's energies for config 1 (FP32 vectors there, f10.5)
    fx1 = config_fixture(1, casedir)[2]
    assert np.abs(runs[0][1][:3] - fx1["res"][:3, 0]).max() < 5e-5


def test_noalloc_refuses_oversized_requests(casedir):
    from bigstickpublick_amd import System
    case, d, fx = config_fixture(1, casedir)
    s = System(*paths(case, d), case.Z, case.N, case.jz2, upload=True)
    rc, *_, msg = _noalloc(s, 10, 400, "f64")
    assert rc != 0 and "bytes of scratch" in msg, msg
    # a thick-restart request is refused too (it would allocate)
    from bigstickpublick_amd.engine import BS_F32, _Params, load_library
    L = load_library()
    p = _Params(5, 40, 10, 3, BS_F32, 0.0, 0)
    e = np.zeros(5)
    assert L.bs_lanczos_noalloc(s.h, ctypes.byref(p), None, None, e.ctypes.data, None) != 0
    # and a valid call still works afterwards
    rc, a, b, e, it, msg = _noalloc(s, 5, 50, "f32")
    assert rc == 0 and np.all(np.isfinite(e)), msg
    s.close()
