"""The synchronous entry points start after all earlier device work on every
stream: a pivot or wave function that a caller writes on a non-blocking
stream, behind a long kernel, is complete when bs_strength or bs_density
reads it (called here through the C ABI with no host synchronization)."""
import ctypes
import os

import numpy as np
import pytest

from golden_cases import paths, small_fixture

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system(tmp_path_factory):
    from bigstickpublick_amd import System
    d = str(tmp_path_factory.mktemp("streams"))
    case, cd, _ = small_fixture("sd2p2n", d)
    s = System(*paths(case, cd), case.Z, case.N, case.jz2, upload=True)
    yield s
    s.close()


def _late_copy(src):
    """A copy of src written on a non-blocking side stream after a spin
    kernel, with no host synchronization afterwards."""
    dst = torch.zeros_like(src)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        torch.cuda._sleep(20_000_000)
        dst.copy_(src)
    return dst, side


def test_strength_waits_for_a_pivot_on_a_side_stream(system):
    from bigstickpublick_amd.engine import BS_F64, _Params, load_library
    s = system
    src = torch.tensor(np.random.default_rng(31).standard_normal(s.dim), device="cuda")
    want = s.strength(src, nkeep=3, niter=40, polish=False)
    piv, side = _late_copy(src)
    p = _Params(3, 40, 0, 0, BS_F64, 0.0, 0)
    m = min(40, s.dim)
    ritz, stren, alpha, beta = (np.zeros(m) for _ in range(4))
    ev, res = np.zeros(3), np.zeros(3)
    pn, it = ctypes.c_double(), ctypes.c_int()
    rc = load_library().bs_strength(s.h, ctypes.byref(p), ctypes.c_void_p(piv.data_ptr()), ctypes.byref(pn),
                                    ritz.ctypes.data, stren.ctypes.data, alpha.ctypes.data, beta.ctypes.data,
                                    ev.ctypes.data, None, res.ctypes.data, ctypes.byref(it))
    assert rc == 0
    side.synchronize()
    assert abs(pn.value - want.pivot_norm) < 1e-9 * want.pivot_norm
    assert np.abs(ev - want.evals).max() < 1e-9
    assert np.abs(stren - want.strength).max() < 1e-9 * max(1.0, np.abs(want.strength).max())


def test_density_waits_for_a_vector_on_a_side_stream(system):
    from bigstickpublick_amd.engine import BS_F64, load_library
    s = system
    src = torch.tensor(np.random.default_rng(32).standard_normal(s.dim), device="cuda")
    rpp, rnn, rpn, n2 = s.density(src)
    psi, side = _late_copy(src)
    out = [np.zeros_like(rpp), np.zeros_like(rnn), np.zeros_like(rpn)]
    norm2 = ctypes.c_double()
    rc = load_library().bs_density(s.h, ctypes.c_void_p(psi.data_ptr()), BS_F64, out[0].ctypes.data,
                                   out[1].ctypes.data, out[2].ctypes.data, ctypes.byref(norm2))
    assert rc == 0
    side.synchronize()
    assert norm2.value == n2
    for a, b in zip(out, (rpp, rnn, rpn)):
        assert np.array_equal(a, b)


def test_torch_keeps_the_gpu_when_the_library_comes_first():
    """A fresh process that builds and uploads a system before it imports
    torch still has torch's GPU afterwards (the binding lets torch set the
    device up before the library's first HIP call)."""
    import subprocess
    import sys
    script = r'''
import os, sys, tempfile
sys.path.insert(0, %r)
from bigstickpublick_amd import System
from bigstickpublick_amd.inputs import small_case, write_case
c = small_case("sd", 2, 2, 0, seed=77)
d = tempfile.mkdtemp()
write_case(c, d)
s = System(os.path.join(d, "sd.sps"), os.path.join(d, c.name + ".int"), 2, 2, 0, upload=True)
import torch
assert torch.cuda.is_available()
x = torch.ones(s.dim, device="cuda", dtype=torch.float64)
print("ok", float(s.apply(x).sum()))
''' % os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
