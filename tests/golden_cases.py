"""Helpers that rebuild the inputs behind each committed fixture."""
import json
import os

import numpy as np

from bigstickpublick_amd.inputs import config_case, small_case, write_case

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL = ["sd2p2n", "sd1p2n_m1", "sd3p1n_mm2", "sd0p3n_m3", "sd1p0n_m1", "sd1p1n",
         "pf2p1n_m3", "pf2p2n", "sdg1p2n_m1", "pf0p8n"]


def small_fixture(name, workdir):
    with open(os.path.join(GOLD, name + ".json")) as f:
        meta = json.load(f)
    case = small_case(meta["space"], meta["Z"], meta["N"], meta["jz2"], seed=meta["seed"])
    case.name = name
    d = os.path.join(workdir, name)
    write_case(case, d)
    fx = dict(np.load(os.path.join(GOLD, name + ".npz")))
    return case, d, fx


def config_fixture(cfg, workdir):
    case = config_case(cfg)
    d = os.path.join(workdir, "cfg%d" % cfg)
    write_case(case, d)
    fx = dict(np.load(os.path.join(GOLD, "config%d.npz" % cfg)))
    return case, d, fx


def ensemble_fixture(workdir, members):
    """The first `members` Hamiltonians of the sd 4p4n ensemble, written into
    one directory, and the reference's results for them."""
    d = os.path.join(workdir, "cfg2")
    cases = []
    for m in range(members):
        case = config_case(2, m)
        write_case(case, d)
        cases.append(case)
    fx = dict(np.load(os.path.join(GOLD, "config2.npz")))
    return cases, d, fx


def paths(case, d):
    return os.path.join(d, case.space + ".sps"), os.path.join(d, case.name + ".int")


def probe_vectors(D):
    i = np.arange(D, dtype=np.float64)
    return np.stack([np.sin(1.3 * i + 0.1), np.full(D, 1.0 / np.sqrt(D))], axis=1)


def dense_from_fixture(fx):
    D = int(fx["dim"])
    R = np.zeros((D, D))
    i, j, v = fx["ham_i"], fx["ham_j"], fx["ham_v"].astype(np.float64)
    R[i, j] = v
    R[j, i] = v
    return R


# The reference's FP64 thick-restart runs kept as <name>_<option>_fp64.res /
# .lcoef, with the parameters the reference derives from each menu's answers
# (src/blanczosmain.f90:993-1053, :1091-1145): td raises niter to
# max((nkeep + 5) * 2, 10) and keeps min(max(nkeep + 5, 3 nkeep), niter - 5);
# tx raises niter the same way and tests convergence as its code does.
THICK_RUNS = [
    ("pf2p1n_m3", "tf", dict(nkeep=4, niter=24, nthick=10, max_restarts=3)),
    ("pf2p1n_m3", "tc", dict(nkeep=4, niter=30, nthick=12, max_restarts=40, tol=1e-4, test=0)),
    ("sd1p2n_m1", "td", dict(nkeep=4, niter=20, nthick=12, max_restarts=40, tol=1e-3, test=0)),
    ("sd1p2n_m1", "tx", dict(nkeep=4, niter=30, nthick=12, max_restarts=200, tol=1e-3, test=0, etarget=-16.2)),
    ("pf2p1n_m3", "tx", dict(nkeep=4, niter=30, nthick=12, max_restarts=200, tol=1e-3, test=0, etarget=-26.5)),
    # option tc with the Ritz-vector tests 2 and 3
    ("sd2p2n", "tc2", dict(nkeep=5, niter=30, nthick=12, max_restarts=40, tol=1e-4, test=2)),
    ("pf2p1n_m3", "tc3", dict(nkeep=5, niter=30, nthick=12, max_restarts=40, tol=1e-4, test=3)),
]


def read_lcoef(path):
    """Line labels, alpha, beta of a .lcoef (negative labels: the kept Ritz
    values and their couplings after a thick restart)."""
    k, a, b = [], [], []
    for line in open(path):
        p = line.split()
        if len(p) == 3:
            k.append(int(p[0]))
            a.append(float(p[1]))
            b.append(float(p[2]))
    return np.array(k), np.array(a), np.array(b)


def res_iterations(path):
    """(iterations, total or None) from a .res header: the size of the final
    projected matrix and, after thick restarts, iterations + niter * restarts."""
    import re
    text = open(path, errors="replace").read()
    it = int(re.search(r"(\d+)\s+iterations", text).group(1))
    m = re.search(r"total =\s+(\d+)", text)
    return it, (int(m.group(1)) if m else None)
