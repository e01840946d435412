"""Device entry points of the Python binding refuse bad tensors before any
library call: host tensors, wrong sizes, mismatched output buffers.  The CPU
cases need no GPU (the system is never uploaded); the GPU cases check the
size, dtype and overlap rules on real device tensors."""
import numpy as np
import pytest
import torch

from bigstickpublick_amd.engine import System
from golden_cases import paths, small_fixture


@pytest.fixture(scope="module")
def sysm(tmp_path_factory):
    case, d, _ = small_fixture("sd2p2n", str(tmp_path_factory.mktemp("chk")))
    s = System(*paths(case, d), case.Z, case.N, case.jz2)
    yield s
    s.close()


def test_host_tensors_refused_before_upload(sysm):
    D = sysm.dim
    x = torch.zeros(D, dtype=torch.float64)
    calls = [
        lambda: sysm.apply(x),
        lambda: sysm.observable("J2", x),
        lambda: sysm.density(x),
        lambda: sysm.gradient(x),
        lambda: sysm.strength(x, 2, 10),
        lambda: sysm.block_strength(x.reshape(1, D), 3),
        lambda: sysm.operator_matrix(None, x.reshape(1, D)),
        lambda: sysm.apply_onebody("none.opme", x),
        lambda: sysm.apply_block(x.reshape(D, 1)),
        lambda: sysm.write_wfn("unused.wfn", [0.0], [0.0], [0.0], x.reshape(1, D).float().to("meta")),
    ]
    for c in calls:
        with pytest.raises(ValueError):
            c()
    # nothing above reached the library's device side
    assert not sysm.uploaded


def test_non_tensors_refused(sysm):
    with pytest.raises(ValueError):
        sysm.apply(np.zeros(sysm.dim))
    with pytest.raises(ValueError):
        sysm.block_strength(np.zeros((1, sysm.dim)), 3)
    assert not sysm.uploaded


@pytest.mark.gpu
def test_device_tensor_rules(sysm):
    D = sysm.dim
    x = torch.randn(D, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        sysm.apply(x[: D - 1])                                   # short x
    with pytest.raises(ValueError):
        sysm.apply(x, y=torch.empty(D, dtype=torch.float32, device="cuda"))  # dtype mismatch
    with pytest.raises(ValueError):
        sysm.apply(x, y=torch.empty(D - 1, dtype=torch.float64, device="cuda"))  # short y
    with pytest.raises(ValueError):
        sysm.apply(x, y=torch.empty(2 * D, dtype=torch.float64, device="cuda")[::2])  # strided y
    with pytest.raises(ValueError):
        sysm.apply(x, y=x)                                       # in place
    with pytest.raises(ValueError):
        sysm.apply(x.to(torch.int64))                             # integer vector
    with pytest.raises(ValueError):
        sysm.gradient(x[:10])
    with pytest.raises(ValueError):
        sysm.apply_block(torch.randn(D, 65, dtype=torch.float64, device="cuda"))  # K > 64
    # a correct call still works after the refusals
    y = sysm.apply(x)
    assert torch.isfinite(y).all()
