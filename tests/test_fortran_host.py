"""The Fortran host (bigstickpublick_amd/fortran/bs_driver, iso_c_binding to the
C ABI) driven by the reference's own menu answers."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bigstickpublick_amd.build import FDRIVER
from bigstickpublick_amd.inputs import menu_script
from golden_cases import config_fixture

needs_driver = pytest.mark.skipif(not os.path.exists(FDRIVER), reason="Fortran driver not built (no amdflang)")


def _run(script, cwd, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([FDRIVER], input=script, text=True, capture_output=True, cwd=cwd,
                          env=env, timeout=600)


@needs_driver
def test_fortran_host_builds_tables(casedir):
    case, d, fx = config_fixture(0, casedir)
    r = _run(menu_script(case, "n", "lf"), d, {"BS_HOST_ONLY": "1"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dimension = %d" % int(fx["dim"]) in r.stdout


@needs_driver
def test_fortran_host_rejects_bad_menu(casedir):
    case, d, fx = config_fixture(0, casedir)
    r = _run("d\n" + menu_script(case, "n", "lf"), d)
    assert r.returncode != 0


def read_res(path):
    rows = []
    lines = open(path).read().splitlines()
    for i, l in enumerate(lines):
        if l.strip().startswith("State"):
            for l2 in lines[i + 1:]:
                p = l2.split()
                if len(p) >= 5 and p[0].isdigit():
                    rows.append([float(p[1]), float(p[3]), float(p[4])])
    return np.array(rows)


@needs_driver
@pytest.mark.gpu
def test_fortran_host_config0_matches_reference(casedir):
    """Same menu script as the reference run: energies, J and T of the .res
    table agree with the reference's .res (f10.5 / f8.3 prints)."""
    case, d, fx = config_fixture(0, casedir)
    wd = os.path.join(casedir, "fdrv0")
    shutil.copytree(d, wd, dirs_exist_ok=True)
    r = _run(menu_script(case, "n", "lf"), wd)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_res(os.path.join(wd, case.name + ".res"))
    ref = fx["res"]
    assert got.shape == ref.shape
    assert np.abs(got[:, 0] - ref[:, 0]).max() < 2e-5
    assert np.abs(got[:, 1] - ref[:, 1]).max() < 2e-3
    assert np.abs(got[:, 2] - ref[:, 2]).max() < 2e-3
    lc = np.loadtxt(os.path.join(wd, case.name + ".lcoef"))
    assert np.abs(lc[:20, 1] - fx["alpha"][:20]).max() < 2e-4
