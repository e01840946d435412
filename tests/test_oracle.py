"""The oracle (oracle/bsoracle.c) pinned against the reference's own outputs
(tests/golden, written by tests/golden/make_golden.py from the reference run)
and against closed forms."""
import math
import os

import numpy as np
import pytest

from golden_cases import (GOLD, SMALL, THICK_RUNS, config_fixture, ensemble_fixture, dense_from_fixture, paths,
                          probe_vectors, read_lcoef, res_iterations, small_fixture)
from lanczos_ref import lanczos
from oracle_api import Oracle, lib


def cg_exact(j1, m1, j2, m2, J, M):
    """Racah's closed formula in exact rationals (arguments are 2j, 2m)."""
    from fractions import Fraction
    if m1 + m2 != M or J < abs(j1 - j2) or J > j1 + j2 or (j1 + j2 + J) % 2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(M) > J:
        return 0.0
    f = math.factorial
    a, b, c = j1 // 1, j2, J

    def h(x):
        return x // 2
    pre = Fraction((J + 1) * f(h(j1 + j2 - J)) * f(h(j1 - j2 + J)) * f(h(-j1 + j2 + J)),
                   f(h(j1 + j2 + J) + 1))
    pre *= f(h(J + M)) * f(h(J - M)) * f(h(j1 - m1)) * f(h(j1 + m1)) * f(h(j2 - m2)) * f(h(j2 + m2))
    s = Fraction(0)
    for k in range(0, 100):
        d = [k, h(j1 + j2 - J) - k, h(j1 - m1) - k, h(j2 + m2) - k, h(J - j2 + m1) + k,
             h(J - j1 - m2) + k]
        if min(d) < 0:
            if d[1] < 0 or d[2] < 0 or d[3] < 0:
                break
            continue
        den = 1
        for x in d:
            den *= f(x)
        s += Fraction((-1) ** k, den)
    del a, b, c
    return float(s) * math.sqrt(float(pre))


def test_cleb_against_racah_formula():
    """src/libra.f CLEB (restated in float) against exact Racah values."""
    L = lib()
    worst = 0.0
    for j1 in range(1, 10, 2):
        for j2 in range(1, 10, 2):
            for J in range(abs(j1 - j2), j1 + j2 + 1, 2):
                for m1 in range(-j1, j1 + 1, 2):
                    for m2 in range(-j2, j2 + 1, 2):
                        M = m1 + m2
                        if abs(M) > J:
                            continue
                        got = L.bso_cleb(j1, m1, j2, m2, J, M)
                        worst = max(worst, abs(got - cg_exact(j1, m1, j2, m2, J, M)))
    assert worst < 2e-6


def test_cleb_special_cases():
    L = lib()
    for j in range(1, 12, 2):
        for m in range(-j, j + 1, 2):
            want = (-1) ** ((j - m) // 2) / math.sqrt(j + 1)
            assert abs(L.bso_cleb(j, m, j, -m, 0, 0) - want) < 1e-6
        assert abs(L.bso_cleb(j, j, j, j, 2 * j, 2 * j) - 1.0) < 1e-6


@pytest.mark.parametrize("name", SMALL)
def test_basis_matches_reference(name, casedir):
    case, d, fx = small_fixture(name, casedir)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    assert o.dim == int(fx["dim"])
    mp, sp, _ = o.sds(0)
    mn, sn, _ = o.sds(1)
    np.testing.assert_array_equal(mp, fx["pmask"])
    np.testing.assert_array_equal(mn, fx["nmask"])
    st = fx["states"]
    np.testing.assert_array_equal(sp[st[:, 1]] + sn[st[:, 2]], st[:, 0])


@pytest.mark.parametrize("name", SMALL)
def test_hamiltonian_matches_reference(name, casedir):
    """H from the oracle against the reference's `wh` dump (ham.dat): the
    reference sums rows in FP32, so agreement is to a few FP32 ulp."""
    case, d, fx = small_fixture(name, casedir)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    if "ham_hx" in fx:
        X = probe_vectors(o.dim)
        for k in range(X.shape[1]):
            y = o.apply(X[:, k])
            ref = fx["ham_hx"][:, k]
            assert np.abs(y - ref).max() < 4e-6 * max(1.0, np.abs(ref).max())
        return
    H = o.dense()
    R = dense_from_fixture(fx)
    scale = np.maximum(np.abs(R), 1.0)
    assert (np.abs(H - R) / scale).max() < 4 * 2.0 ** -23 * 16
    assert np.array_equal(np.abs(H) > 1e-5, np.abs(R) > 1e-5)
    # eigenvalues of the two matrices
    e1 = np.linalg.eigvalsh(H)[:5]
    e2 = np.linalg.eigvalsh(R)[:5]
    assert np.abs(e1 - e2).max() < 2e-5


@pytest.mark.parametrize("cfg", [0, 1])
def test_config_basis_and_spectrum(cfg, casedir):
    """Configs 0/1: basis bit-exact; FP64 Lanczos from the same uniform pivot
    reproduces the reference's alpha/beta (FP32 vectors there) and energies."""
    case, d, fx = config_fixture(cfg, casedir)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    assert o.dim == int(fx["dim"])
    mp, sp, _ = o.sds(0)
    mn, sn, _ = o.sds(1)
    np.testing.assert_array_equal(mp, fx["pmask"])
    np.testing.assert_array_equal(mn, fx["nmask"])
    np.testing.assert_array_equal(sp, fx["pstart"])
    np.testing.assert_array_equal(sn, fx["nstart"])
    if cfg != 0:
        return
    a, b, e, _ = lanczos(o.apply, o.dim, 150, case.nkeep)
    ra, rb = fx["alpha"], fx["beta"]
    assert np.abs(a[:30] - ra[:30]).max() < 2e-5
    assert np.abs(b[:30] - rb[:30]).max() < 2e-5
    assert np.abs(e - fx["res"][: case.nkeep, 0]).max() < 2e-5


@pytest.mark.parametrize("member", [3, 6])
def test_ensemble_member_spectrum(member, casedir):
    """Batched-ensemble Hamiltonians (sd 4p4n, SPEs ~ N(0, 2 MeV)): FP64
    Lanczos on the oracle reproduces the reference's alpha/beta; the exact
    lowest eigenvalues (ARPACK on the oracle's H) match every reference state
    that its 150 FP32 steps converged (integer J)."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    cases, d, fx = ensemble_fixture(casedir, member + 1)
    case = cases[member]
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    assert o.dim == 28503
    # the reference's FP32 vectors drift from FP64 by ~1e-6 per step
    a, b, e, _ = lanczos(o.apply, o.dim, 30, case.nkeep)
    assert np.abs(a[:20] - fx["alpha"][member, :20]).max() < 2e-5
    assert np.abs(b[:20] - fx["beta"][member, :20]).max() < 2e-5
    assert np.abs(a - fx["alpha"][member, :30]).max() < 2e-4
    op = LinearOperator((o.dim, o.dim), matvec=lambda v: o.apply(np.ascontiguousarray(v, dtype=np.float64)),
                        dtype=np.float64)
    ex = np.sort(eigsh(op, k=7, which="SA", tol=1e-10, v0=np.ones(o.dim))[0])
    ref = fx["res"][member]
    conv = np.abs(ref[:, 1] - np.round(ref[:, 1])) < 5e-4  # J printed as an integer (f8.3)
    assert conv.sum() >= 3
    for (E, J, T), ok in zip(ref, conv):
        if ok:
            assert np.abs(ex - E).min() < 2e-5, (E, J, ex)
    assert abs(ex[0] - ref[0, 0]) < 2e-5


def test_hermitian_and_planted_mistake(casedir):
    """A flipped sign in one coupled TBME must break agreement with ham.dat."""
    case, d, fx = small_fixture("sd2p2n", casedir)
    sps, intf = paths(case, d)
    o = Oracle(sps, intf, case.Z, case.N, case.jz2)
    H = o.dense()
    assert np.abs(H - H.T).max() == 0.0
    lines = open(intf).read().splitlines()
    p = lines[5].split()
    p[6] = repr(-float(p[6]))
    lines[5] = " ".join(p)
    bad = intf.replace(".int", "_bad.int")
    open(bad, "w").write("\n".join(lines) + "\n")
    Hb = Oracle(sps, bad, case.Z, case.N, case.jz2).dense()
    R = dense_from_fixture(fx)
    assert np.abs(Hb - R).max() > 1e-2


@pytest.mark.parametrize("name", ["sd2p2n", "sd1p0n_m1", "sd0p3n_m3", "pf2p1n_m3"])
def test_factorized_port_matches_bruteforce(name, casedir):
    """The factorized CPU port (reference loop structure, FP32) against the
    brute-force FP64 oracle: two independent constructions of H x."""
    case, d, fx = small_fixture(name, casedir)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    p = o.port()
    x = np.random.default_rng(0).standard_normal(o.dim)
    y, n = p.apply(x)
    want = o.apply(x)
    assert np.abs(y - want).max() < 1e-6 * np.abs(want).max()
    assert n == p.count()


def _mg25_energies(path):
    rows = []
    lines = open(path, errors="replace").read().splitlines()
    for i, l in enumerate(lines):
        if l.strip().startswith("State") and "Ex" in l:
            for l2 in lines[i + 1:]:
                p = l2.split()
                if len(p) < 5 or not p[0].isdigit():
                    break
                rows.append([float(p[1]), float(p[3]), float(p[4])])
            break
    return np.array(rows)


def test_oracle_mg25_usdb_matches_reference_example():
    """The reference's own example (examples/sd: 25Mg, USDB, scaling
    (18/25)^0.3, 2Jz = 1, D = 44,133): the oracle's Hamiltonian, diagonalized
    by the FP64 Lanczos restatement, gives the ten energies of the
    reference's committed mg25ref.res (printed f10.5)."""
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mg25")
    vs = float(np.float32((18.0 / 25.0) ** 0.3))
    o = Oracle(os.path.join(d, "sd.sps"), os.path.join(d, "usdb.int"), 4, 5, 1, 1.0, vs)
    assert o.dim == 44133
    ref = _mg25_energies(os.path.join(d, "mg25ref.res"))
    _, _, e, _ = lanczos(o.apply, o.dim, 220, 10)
    assert ref.shape == (10, 3)
    assert np.abs(e - ref[:, 0]).max() < 6e-6, (e, ref[:, 0])


def test_autoscaled_int_matches_reference_run(tmp_path):
    """An autoscaled .int (negative count; core 16, A0 18, x 0.3 after the
    SPEs): the engine's tables equal, bit for bit, those of the same file
    read with the scaling line (1, 18, 16 + Z + N, 0.3), and the oracle with
    that scale reproduces the reference's own run on the autoscaled file
    (alpha of the first 20 Lanczos steps, energies)."""
    from bigstickpublick_amd.engine import System
    from bigstickpublick_amd.inputs import small_case, write_case
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    c = small_case("sd", 2, 2, 0, seed=77)
    c.name = "sd2p2n"
    d = str(tmp_path)
    write_case(c, d)
    sps = os.path.join(d, "sd.sps")
    auto = System(sps, os.path.join(gold, "sd2p2n_auto.int"), 2, 2, 0)
    plain = System(sps, os.path.join(d, "sd2p2n.int"), 2, 2, 0, scaling=(1.0, 18.0, 20.0, 0.3))
    for k in range(3):
        assert np.array_equal(auto.me_table(k).view(np.uint32), plain.me_table(k).view(np.uint32))
    vs = float(np.float32(18.0) / np.float32(20.0)) ** np.float32(0.3)
    o = Oracle(sps, os.path.join(d, "sd2p2n.int"), 2, 2, 0, 1.0, float(np.float32(vs)))
    a, _, e, _ = lanczos(o.apply, o.dim, 60, 5)
    lc = np.loadtxt(os.path.join(gold, "sd2p2n_auto.lcoef"))
    assert np.abs(a[:20] - lc[:20, 1]).max() < 1e-4
    res = []
    for l in open(os.path.join(gold, "sd2p2n_auto.res")).read().splitlines():
        p = l.split()
        if len(p) == 6 and p[0].isdigit():
            res.append((float(p[1]), float(p[3])))
    res = np.array(res[:5])
    # the states the 60-step runs have converged (integer J in the
    # reference's table); the others still mix and follow the rounding
    conv = np.abs(res[:, 1] - np.round(res[:, 1])) < 1e-3
    assert conv.sum() >= 3
    assert np.abs(e - res[:, 0])[conv].max() < 2e-5, (e, res)


def test_oracle_row_apply_equals_full_apply(casedir):
    """bso_apply_rows (the row-sampled oracle used at sizes where the full
    apply is too slow) gives, on every listed proton row, the very bits of
    the full apply, and leaves the other rows zero."""
    case, d, fx = small_fixture("pf2p2n", casedir)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    x = np.random.default_rng(4).standard_normal(o.dim)
    full = o.apply(x)
    _, pstart, _ = o.sds(0)
    nP = len(pstart)
    rows = np.array([0, 3, nP // 2, nP - 1])
    part = o.apply_rows(x, rows)
    bounds = np.append(np.sort(pstart), o.dim)
    mask = np.zeros(o.dim, bool)
    for r in rows:
        lo = pstart[r]
        hi = bounds[np.searchsorted(bounds, lo, side="right")]
        mask[lo:hi] = True
    assert mask.sum() > 0
    assert np.array_equal(part[mask], full[mask])
    assert not np.any(part[~mask])


def test_convergence_rule_reproduces_reference_stopping_points():
    """The restated default convergence test (lanczos_ref.convergence_stop),
    replayed on the reference's own coefficients of its option-ld runs, stops
    exactly where the reference's .lcoef ends: sd 2p2n after 100 iterations
    (mean change 7.4e-4 < 1e-3 there, 0.37 at 90), config 0 never before its
    150-iteration cap."""
    from lanczos_ref import convergence_stop
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    a, b = _read_lcoef(os.path.join(gold, "sd2p2n_ld.lcoef"))
    assert len(a) == 100
    assert convergence_stop(a, b, 5, niter=300) == 100
    # planted variants of the rule do not stop there: the largest change
    # instead of the mean, no extra states (stops at 80), a tighter tolerance
    assert convergence_stop(a, b, 5, niter=300, test=1) != 100
    assert convergence_stop(a, b, 5, ncheck_ex=0, niter=300) == 80
    assert convergence_stop(a, b, 5, tol=3e-4, niter=300) != 100
    a0, b0 = _read_lcoef(os.path.join(gold, "sd4p4n_ld.lcoef"))
    assert len(a0) == 150 and convergence_stop(a0, b0, 5, niter=150) == 150
    for name, stop, cap in (("sd1p2n_m1", 70, 127), ("pf2p1n_m3", 100, 300)):
        a1, b1 = _read_lcoef(os.path.join(gold, "%s_ld.lcoef" % name))
        assert len(a1) == stop and convergence_stop(a1, b1, 5, niter=cap) == stop, name


def test_wavefunction_convergence_rule_reproduces_reference_stopping_points():
    """Tests 2 and 3 (the Ritz vectors of successive checks compared), the
    restatement replayed on the reference's own coefficients of its option-lc
    runs, stop exactly where each .lcoef ends: sd 2p2n test 2 at 1e-4 after 80
    iterations, test 3 at 1e-4 after 90, pf 2p1n test 3 at 1e-5 after 90; with
    two extra check states (sd 1p2n) the reference's plain run has no vectors
    for them (only the lowest nkeep are solved for), so the test never passes
    and the run goes to its cap."""
    from lanczos_ref import convergence_stop
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for name, test, tol, nex, stop in (("sd2p2n_lc2", 2, 1e-4, 0, 80), ("sd2p2n_lc3", 3, 1e-4, 0, 90),
                                       ("pf2p1n_m3_lc3", 3, 1e-5, 0, 90), ("sd1p2n_m1_lc2x", 2, 1e-4, 2, 128)):
        a, b = _read_lcoef(os.path.join(gold, name + ".lcoef"))
        assert len(a) == stop, name
        assert convergence_stop(a, b, 5, ncheck_ex=nex, tol=tol, test=test, niter=max(stop, 300) if nex == 0
                                else stop) == stop, name
    # planted variants do not stop there: the other wave-function test, a
    # tighter tolerance, vectors for the extra states
    a, b = _read_lcoef(os.path.join(gold, "sd2p2n_lc2.lcoef"))
    a3, b3 = _read_lcoef(os.path.join(gold, "sd2p2n_lc3.lcoef"))
    assert convergence_stop(a3, b3, 5, ncheck_ex=0, tol=1e-4, test=2, niter=300) == 80
    assert convergence_stop(a3, b3, 5, ncheck_ex=0, tol=1e-4, test=3, niter=300) == 90
    assert convergence_stop(a3, b3, 5, ncheck_ex=0, tol=1e-7, test=2, niter=300) != 80
    ax, bx = _read_lcoef(os.path.join(gold, "sd1p2n_m1_lc2x.lcoef"))
    assert convergence_stop(ax, bx, 7, ncheck_ex=0, tol=1e-4, test=2, niter=128) < 128


def _read_lcoef(path):
    a, b = [], []
    for line in open(path):
        p = line.split()
        if len(p) == 3:
            a.append(float(p[1]))
            b.append(float(p[2]))
    return np.array(a), np.array(b)


def test_oracle_lanczos_equals_reference_fp64_build(casedir):
    """The FP64 Lanczos restatement on the oracle's Hamiltonian against the
    reference's own FP64 build (oracle/Makefile ref64, lanc_prec = 8; its
    config-0 .lcoef and .res in tests/golden/config0_fp64.npz): the
    coefficients agree to 1e-10 for 60 steps and the five Ritz values of the
    150-step run to 1e-9 MeV (measured 4e-12 and 5e-13)."""
    case, d, fx = config_fixture(0, casedir)
    f64 = dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config0_fp64.npz")))
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    a, b, e, _ = lanczos(o.apply, o.dim, 150, 5)
    assert np.abs(a[:60] - f64["alpha"][:60]).max() < 1e-10
    assert np.abs(b[:60] - f64["beta"][:60]).max() < 1e-10
    al, be = f64["alpha"], f64["beta"]
    ref = np.linalg.eigvalsh(np.diag(al) + np.diag(be[:-1], 1) + np.diag(be[:-1], -1))[:5]
    assert np.abs(e - ref).max() < 1e-9
    # the printed energies (f10.5) are these Ritz values
    assert np.abs(f64["res"][:5, 0] - ref).max() < 6e-6


def _res_energies(path):
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


@pytest.mark.parametrize("name,option,kw", THICK_RUNS, ids=["%s-%s" % (r[0], r[1]) for r in THICK_RUNS])
def test_oracle_thick_restart_equals_reference_fp64_build(name, option, kw, casedir):
    """The thick-restart restatement (lanczos_ref.thick_lanczos) against the
    reference's FP64 build on the same menu answers (tf, tc, td, tx): the same
    final matrix size and number of restarts (.res "iterations" and "total"),
    the .lcoef -- kept Ritz values, their couplings up to sign (an
    eigenvector's sign is free), then the last cycle's steps -- to 1e-4, the
    .res energies to 2e-5.  The coefficients agree to 1e-8 .. 2e-5, not to
    1e-12: on these cases the Hamiltonians already differ at the reference's
    FP32 input rounding (the first alpha by 4e-8) and unconverged Ritz values
    amplify it; a cycle one vector longer moves them by about 1.  The tx runs
    stop at the first check (iteration 10): the reference's tx tests no
    energies."""
    from lanczos_ref import thick_lanczos
    case, d, _ = small_fixture(name, casedir)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    r = thick_lanczos(o.apply, o.dim, **kw)
    base = os.path.join(GOLD, "%s_%s_fp64" % (name, option))
    it, total = res_iterations(base + ".res")
    assert r["steps"] == it
    assert (total if total is not None else it) == it + kw["niter"] * r["restarts"]
    k, a, b = read_lcoef(base + ".lcoef")
    assert len(a) == it
    assert (k < 0).sum() == (kw["nthick"] if r["restarts"] else 0)
    assert np.abs(a - r["alpha"]).max() < 1e-4
    assert np.abs(np.abs(b) - np.abs(r["beta"])).max() < 1e-4
    assert np.abs(_res_energies(base + ".res") - r["evals"]).max() < 2e-5
    if option == "tx":
        assert it == 10 and r["restarts"] == 0


def test_oracle_targeted_run_converges_to_states_nearest_target(casedir):
    """tx with the test over the nkeep kept energies (no reference run: the
    reference's tx never tests them): the four converged energies are the
    four eigenvalues of the dense H nearest the target, sorted."""
    from lanczos_ref import thick_lanczos
    for name, et in (("sd1p2n_m1", -16.2), ("pf2p1n_m3", -26.5)):
        case, d, _ = small_fixture(name, casedir)
        o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
        ev = np.linalg.eigvalsh(o.dense())
        want = np.sort(ev[np.argsort(np.abs(ev - et), kind="stable")[:4]])
        r = thick_lanczos(o.apply, o.dim, nkeep=4, niter=30, nthick=12, max_restarts=400, tol=1e-9, test=1,
                          etarget=et, target_test="kept")
        assert r["restarts"] <= 400
        assert np.abs(r["evals"] - want).max() < 1e-6, (name, r["evals"], want)


def test_oracle_long_rows_equals_reference_run(casedir):
    """pf 1p8n (2Jz = 1, D = 219,781; neutron sectors up to 12,022 SDs): the
    FP64 Lanczos restatement on the oracle's H against the reference's lf
    run: 20 coefficients agree to its FP32-vector rounding (2e-5)."""
    case, d, fx = small_fixture("pf1p8n_m1", casedir)
    o = Oracle(*paths(case, d), case.Z, case.N, case.jz2)
    assert o.dim == 219781
    a, b, _, _ = lanczos(o.apply, o.dim, 20, 5)
    assert np.abs(a - fx["alpha"][:20]).max() < 2e-5
    assert np.abs(b - fx["beta"][:20]).max() < 2e-5
