"""Thick-restart Lanczos (options td, tf, tc, tx) through the C ABI against the
oracle's restatement of the reference's cycle and the reference's FP64 runs.

The cycle (lanczos_p, src/blanczosmain.f90:1690-1752; thick_restart_sub_p,
src/blanczoslib.f90:1656-1799): steps to niter - 1, restart from the matrix on
the first niter - 2 vectors, v_{niter-1} the next start vector, tests at every
itskipeig-th step, max_restarts + 1 restarts at most.  The FP64 runs equal the
oracle's to 1e-7 (same Hamiltonian; summation order only), and so meet the
reference's fixtures where the oracle does."""
import os

import numpy as np
import pytest

from golden_cases import GOLD, THICK_RUNS, paths, read_lcoef, res_iterations, small_fixture
from lanczos_ref import thick_lanczos
from oracle_api import Oracle

pytestmark = pytest.mark.gpu

IDS = ["%s-%s" % (r[0], r[1]) for r in THICK_RUNS]


def _kwargs(kw, **extra):
    a = dict(nkeep=kw["nkeep"], niter=kw["niter"], thick=kw["nthick"], max_restarts=kw["max_restarts"],
             tol=kw.get("tol", 0.0), conv_test=kw.get("test", 0), target=kw.get("etarget"))
    a.update(extra)
    return a


@pytest.fixture(scope="module")
def systems(tmp_path_factory):
    from bigstickpublick_amd import System
    d = str(tmp_path_factory.mktemp("thick"))
    out = {}
    for name in sorted({r[0] for r in THICK_RUNS}):
        case, cd, _ = small_fixture(name, d)
        out[name] = (System(*paths(case, cd), case.Z, case.N, case.jz2, upload=True),
                     Oracle(*paths(case, cd), case.Z, case.N, case.jz2))
    yield out
    for s, _ in out.values():
        s.close()


@pytest.mark.parametrize("name,option,kw", THICK_RUNS, ids=IDS)
def test_thick_restart_fp64_equals_oracle_and_reference(name, option, kw, systems):
    s, o = systems[name]
    want = thick_lanczos(o.apply, o.dim, **kw)
    r = s.lanczos(dtype="f64", polish=False, **_kwargs(kw))
    _, _, steps, restarts = s.lanczos_stats()
    print("%s %s: steps %d restarts %d (oracle %d %d), E %s" % (name, option, steps, restarts, want["steps"],
                                                                want["restarts"], r.evals))
    assert (steps, restarts) == (want["steps"], want["restarts"])
    assert np.abs(r.alpha[:steps] - want["alpha"]).max() < 1e-7
    assert np.abs(np.abs(r.beta[:steps]) - np.abs(want["beta"])).max() < 1e-7
    assert np.abs(r.evals - want["evals"]).max() < 1e-8
    # the reference's FP64 run of the same menu
    base = os.path.join(GOLD, "%s_%s_fp64" % (name, option))
    it, total = res_iterations(base + ".res")
    assert steps == it and (total if total is not None else it) == it + kw["niter"] * restarts
    _, a, b = read_lcoef(base + ".lcoef")
    assert np.abs(a - r.alpha[:steps]).max() < 1e-4
    assert np.abs(np.abs(b) - np.abs(r.beta[:steps])).max() < 1e-4


@pytest.mark.parametrize("name,option,kw", THICK_RUNS, ids=IDS)
def test_thick_restart_mixed_precision(name, option, kw, systems):
    """FP32 vectors, FP64 coefficients: the converged runs (tc, td) reach the
    oracle's energies within their tolerance's reach (2e-4) and residuals
    below 1e-2; tf and tx (fixed restarts, first-check stop) take the
    oracle's number of steps and restarts, and their Ritz values agree to
    1e-4."""
    s, o = systems[name]
    want = thick_lanczos(o.apply, o.dim, **kw)
    r = s.lanczos(dtype="f32", polish=True, **_kwargs(kw))
    _, _, steps, restarts = s.lanczos_stats()
    print("%s %s mixed: steps %d restarts %d (oracle %d %d), E %s, resid %s" % (
        name, option, steps, restarts, want["steps"], want["restarts"], r.evals, r.resid))
    if option in ("tf", "tx"):
        assert (steps, restarts) == (want["steps"], want["restarts"])
        r2 = s.lanczos(dtype="f32", polish=False, **_kwargs(kw))
        assert np.abs(r2.evals - want["evals"]).max() < 1e-4
    else:
        assert np.abs(r.evals - want["evals"]).max() < 2e-4
        assert r.resid.max() < 1e-2


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_targeted_run_converges_to_nearest_states(dtype, systems):
    """target_test="kept": the convergence test over the nkeep kept energies;
    the four energies reached are the dense spectrum's four nearest the
    target (interior states), sorted."""
    for name, et in (("sd1p2n_m1", -16.2), ("pf2p1n_m3", -26.5)):
        s, o = systems[name]
        ev = np.linalg.eigvalsh(o.dense())
        want = np.sort(ev[np.argsort(np.abs(ev - et), kind="stable")[:4]])
        tol = 1e-9 if dtype == "f64" else 1e-6
        r = s.lanczos(nkeep=4, niter=30, thick=12, max_restarts=400, dtype=dtype, tol=tol, conv_test=1,
                      target=et, target_test="kept", polish=True)
        print(name, dtype, r.evals, want, r.resid)
        assert np.abs(r.evals - want).max() < (1e-6 if dtype == "f64" else 2e-4), (name, r.evals, want)


def test_target_needs_thick_restart(systems):
    s, _ = systems["sd1p2n_m1"]
    with pytest.raises(RuntimeError, match="thick"):
        s.lanczos(nkeep=4, niter=30, thick=0, dtype="f64", target=-16.0)
    with pytest.raises(RuntimeError, match="niter - 2"):
        s.lanczos(nkeep=4, niter=30, thick=29, max_restarts=2, dtype="f64")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_block_thick_restart(dtype, systems):
    """Option tb (thick-restart block Lanczos, blocks of 4, restart when the
    basis would reach 20 vectors, 8 kept, the reference's test 0 at 1e-3 over
    nkeep + 2 energies): restarts happen and the run meets the dense
    spectrum to its test's reach, as the reference's own tb run does (its
    random pivot block differs from this seeded one, so the runs agree to that
    reach, not digit for digit); run to a tight test, the restarted block
    Krylov space gives the dense eigenvalues themselves."""
    s, o = systems["pf2p1n_m3"]
    ev = np.linalg.eigvalsh(o.dense())
    ref = _res(os.path.join(GOLD, "pf2p1n_m3_tb_fp64.res"))
    r = s.block_lanczos(nkeep=4, niter=20, block=4, dtype=dtype, tol=1e-3, thick=8, max_restarts=50,
                        conv_test=0, ncheck_ex=2, polish=True)
    _, _, steps, restarts = s.lanczos_stats()
    print("tb %s: %d matvecs, %d restarts, E %s, ref %s, exact %s" % (dtype, r.matvecs, restarts, r.evals, ref,
                                                                     ev[:4]))
    assert restarts >= 1
    assert np.abs(r.evals - ev[:4]).max() < 5e-4
    assert np.abs(r.evals - ref).max() < 7e-4
    t = s.block_lanczos(nkeep=4, niter=20, block=4, dtype=dtype, tol=1e-10 if dtype == "f64" else 1e-6, thick=8,
                        max_restarts=400, conv_test=1, ncheck_ex=2, polish=True)
    assert np.abs(t.evals - ev[:4]).max() < (1e-8 if dtype == "f64" else 2e-4), (t.evals, ev[:4])


def _res(path):
    rows = []
    lines = open(path, errors="replace").read().splitlines()
    for i, line in enumerate(lines):
        if line.strip().startswith("State") and "Ex" in line:
            for l2 in lines[i + 1:]:
                p = l2.split()
                if len(p) < 5 or not p[0].isdigit():
                    break
                rows.append(float(p[1]))
            break
    return np.array(rows)


@pytest.mark.parametrize("name,label,test,tol,nex,stop,exact", [
    ("sd2p2n", "lc2", 2, 1e-4, 0, 80, False), ("sd2p2n", "lc3", 3, 1e-4, 0, 90, True),
    ("pf2p1n_m3", "lc3", 3, 1e-5, 0, 90, True), ("sd1p2n_m1", "lc2x", 2, 1e-4, 2, 128, True)])
def test_ritz_vector_convergence_tests_stop_where_reference_does(name, label, test, tol, nex, stop, exact,
                                                                 tmp_path):
    """Plain runs with the Ritz-vector tests (option lc, tests 2 and 3): each
    run (FP64 and mixed) stops where the restated rule stops on its own
    coefficients, and that is where the reference's own run ended (its .lcoef
    length) unless rounding decides it: on sd 2p2n three low states enter the
    Krylov space of the uniform pivot only through rounding (as with option
    ld), and at test 2's 1e-4 the FP64 run's vectors still move at iteration
    80 (it stops at 90) where the reference's FP32 run's had settled.  With
    extra check states (sd 1p2n) the reference has no vectors for them and
    runs to the cap, and so does this one."""
    from bigstickpublick_amd import System
    from lanczos_ref import convergence_stop
    case, cd, _ = small_fixture(name, str(tmp_path))
    s = System(*paths(case, cd), case.Z, case.N, case.jz2, upload=True)
    ref = read_lcoef(os.path.join(GOLD, "%s_%s.lcoef" % (name, label)))[1]
    assert len(ref) == stop
    cap = min(300, s.dim)
    for dtype in ("f64", "f32"):
        r = s.lanczos(nkeep=5, niter=300, dtype=dtype, tol=tol, conv_test=test, ncheck_ex=nex, polish=False)
        steps = s.lanczos_stats()[2]
        print(name, label, dtype, "steps", steps, "reference", stop)
        assert convergence_stop(r.alpha[:steps], r.beta[:steps], 5, ncheck_ex=nex, tol=tol, test=test,
                                niter=cap) == steps, dtype
        if exact:
            assert steps == stop, (dtype, steps)
        else:
            assert stop <= steps <= stop + 10, (dtype, steps)
    with pytest.raises(RuntimeError, match="target"):
        s.lanczos(nkeep=4, niter=30, thick=12, max_restarts=2, dtype="f64", tol=1e-4, conv_test=2, target=-20.0)
    s.close()
