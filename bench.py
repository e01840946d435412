"""Benchmark: factorized M-scheme H x on one MI355X.

Workload: pf valence space (0f7/2 1p3/2 0f5/2 1p1/2), 6 protons + 6 neutrons,
M = 0, D = 109,954,620, seeded Gaussian two-body ensemble; FP32 vectors and
FP32 matrix elements.  One step = one full H x over the resident vector.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  value = H x per second on the whole job (one GPU).

roofline: the H x kernel's HBM bytes per launch, measured with rocprofv3
(FETCH_SIZE + WRITE_SIZE, scaled by a plain copy kernel of known size;
profiles/k_apply_tiles_pf6p6n_traffic.json, per kernel variant) divided by the
kernel's mean duration timed here with HIP events, against the 8 TB/s HBM
peak.  Beside it, labelled and never substituted: the compulsory bytes (x read
once, y written once) and the stored-matrix equivalent B_alg = N_ops/2 * 8 +
3*4*D bytes (an explicit Hermitian CSR of the same H, FP32 value + int32
index per stored nonzero; N_ops is the reference's own operation count,
src/bigstick_main.f90:1150-1192, which this engine reproduces), which the
factorized apply never reads and which is why its "fraction" exceeds 1.

solve (on by default; --no-solve skips it): the whole eigensolve of the
workload in FP64 and in mixed precision under one stopping rule, eigenpairs
per second within 1e-4 MeV of a tightly converged FP64 run.
"""
import argparse
import dataclasses
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H·v/s + HBM GB/s at pf 6p6n dim 1.1e8; mixed vs FP64 eigenpairs/s at ΔE≤1e-4MeV"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--config", type=int, default=4,
                    help="benchmark configuration (4 = pf 6p6n; 2 = the 64-Hamiltonian sd 4p4n ensemble, batched)")
    ap.add_argument("--solve", dest="solve", action="store_true", default=True,
                    help="time the whole eigensolve (thick-restart Lanczos, keep 10, restart at 40, "
                         "lowest 5, FP64 polish) in FP64 and in mixed precision, and report eigenpairs/s "
                         "at |E - E_FP64| <= 1e-4 MeV (default)")
    ap.add_argument("--no-solve", dest="solve", action="store_false", help="skip the eigensolve leg")
    ap.add_argument("--max-restarts", type=int, default=40)
    ap.add_argument("--batch", type=int, default=0,
                    help="also time K Hamiltonians on the same basis (the workload's interaction and K-1 "
                         "more seeded ones) as one multi-vector matvec; with --solve, the batched "
                         "mixed-precision eigensolve with gradients, scored against FP64")
    ap.add_argument("--block", type=int, default=4,
                    help="also time the workload's H on K interleaved vectors (the block apply of block "
                         "Lanczos) and report H x/s per vector (default K = 4; 0 skips it)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="force this kernel variant (index into System.variants()) instead of autotuning")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="compute HBM bytes per launch of the H x kernel from two rocprofv3 --pmc result "
                         "databases (FETCH_SIZE, WRITE_SIZE; runs made with --copy-probe --variant V), "
                         "write profiles/k_apply_tiles_pf6p6n_traffic.json and exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after timing, copy x into y once (a plain copy kernel of 8*D bytes) "
                         "so hardware-counter runs can calibrate FETCH_SIZE/WRITE_SIZE")
    return ap.parse_args()


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "k_apply_tiles_pf6p6n_traffic.json")


def traffic_from_profiles(variant):
    """HBM bytes per launch of the given kernel variant, measured with
    rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE) and scaled by a copy-kernel
    calibration; committed under profiles/ (one entry per profiled variant)."""
    if not os.path.exists(TRAFFIC_JSON):
        return None
    with open(TRAFFIC_JSON) as f:
        t = json.load(f)
    for e in t.get("variants", [t]):
        if e.get("variant") == list(variant):
            return e
    return None


def traffic_from_dbs(fetch_db, write_db, D):
    """Per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of the H x kernel, scaled by
    what the same counters read for the copy probe (one plain copy of 4*D
    bytes read and 4*D written, the last copy dispatch of the run)."""
    import sqlite3
    out = {}
    for key, db in (("fetch", fetch_db), ("write", write_db)):
        c = sqlite3.connect(db)
        rows = list(c.execute("select dispatch_id, kernel_name, value, duration from counters_collection "
                              "order by dispatch_id"))
        kern = [r for r in rows if "k_apply_tiles<float" in r[1]]
        copies = [r for r in rows if "copy" in r[1].lower() and r[2] > 1e5]
        if not kern or not copies:
            raise RuntimeError("no H x or copy-probe dispatch in " + db)
        names = sorted(set(r[1] for r in kern))
        if len(names) != 1:
            raise RuntimeError("several H x kernels in %s (run with --variant)" % db)
        probe = copies[-1]
        scale = 4.0 * D / (probe[2] * 1024.0)
        vals = [r[2] for r in kern]
        out[key] = {"kib_raw": vals, "probe_kib": probe[2], "probe_kernel": probe[1], "scale": scale,
                    "bytes_per_launch": sum(vals) / len(vals) * 1024.0 * scale,
                    "mean_duration_ns": sum(r[3] for r in kern) / len(kern)}
        out["kernel"] = names[0]
    out["bytes_per_launch"] = out["fetch"]["bytes_per_launch"] + out["write"]["bytes_per_launch"]
    return out


def cpu_baseline(sps, intf, case, target_s=15.0):
    """CPU leg: the factorized FP32 OpenMP port of the reference's apply
    (oracle/bsoracle.c, bso_port_*; same loop structure as
    src/bapplyhlibOMP.f90:979-1450) on this host's cores, timed on a bounded
    sample of whole proton rows of the same workload and scaled to full H x by
    the ratio of updates (the port counts its own updates per row)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    from oracle_api import Oracle
    # the box's CPU share for one GPU is 16 threads whatever nproc reports
    threads = min(len(os.sched_getaffinity(0)), 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = Oracle(sps, intf, case.Z, case.N, case.jz2)
    p = o.port()
    total = p.count()
    x = np.random.default_rng(7).standard_normal(o.dim).astype(np.float32)
    lo = p.nP // 3
    step = max(1, p.nP // 4000)

    def rows_for(work):
        hi, acc = lo, 0
        while hi < p.nP and acc < work:
            nxt = min(p.nP, hi + step)
            acc += p.count(hi, nxt)
            hi = nxt
        return hi

    # calibrate on ~1 s of work (rate guess 1e9 updates/s), then size the sample
    rate = 1e9
    for _ in range(2):
        hi = rows_for(rate * 1.0)
        t0 = time.time()
        _, n = p.apply(x, lo, hi)
        rate = n / max(time.time() - t0, 1e-3)
    hi2 = rows_for(rate * target_s)
    t0 = time.time()
    _, n = p.apply(x, lo, hi2)
    dt = time.time() - t0
    frac = n / total
    return {"value": frac / dt, "unit": "H·v/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "sample": "factorized FP32 OpenMP port of the reference apply (oracle/bsoracle.c) on "
                      "proton rows %d..%d of %d = %.2f%% of the %d updates of one H x, %.2f s; "
                      "scaled by the update fraction" % (lo, hi2, p.nP, 100 * frac, total, dt)}


SOLVE_TOL = 1e-5  # the reference's largest-change test (option tc, test 1) at this tolerance (MeV)


def solve_leg(s, case, max_restarts):
    """Eigenpairs/s of the whole solve (tables resident, Lanczos, restarts,
    polish) for the plain FP64 path and the mixed path (FP32 vectors and
    elements, FP64 alpha/beta, reorthogonalization, Ritz solve and polish),
    both under the same stopping rule (SOLVE_TOL, the reference's thick-restart
    cycle and test schedule).  Each counts the eigenpairs within 1e-4 MeV of
    the scoring energies: per state the lower of two tightly converged runs
    (test at 1e-7 MeV, FP64 polish), one FP64 and one mixed, untimed for the
    ratio.  Both are Ritz values, upper bounds of the true eigenvalues; the
    lower one is the run that has the state -- from the uniform pivot some
    states enter only through rounding, sooner in FP32."""
    import numpy as np
    import torch
    thick, niter, nkeep = case.thick, case.niter, case.nkeep
    out = {"nkeep": nkeep, "niter": niter, "thick": thick, "stop_tol_mev": SOLVE_TOL}
    runs = {}
    for label, dt, tol in (("fp64", "f64", SOLVE_TOL), ("mixed", "f32", SOLVE_TOL), ("exact", "f64", 1e-7),
                           ("exact_mixed", "f32", 1e-7)):
        s.autotune(dt)
        torch.cuda.synchronize()
        t0 = time.time()
        r = s.lanczos(nkeep=nkeep, niter=niter, thick=thick, max_restarts=max_restarts, dtype=dt,
                      tol=tol, conv_test=1, polish=True)
        torch.cuda.synchronize()
        runs[label] = (r, time.time() - t0)
        print("solve %s: %.2f s, %d matvecs" % (label, runs[label][1], r.matvecs), file=sys.stderr, flush=True)
    for label in ("exact", "exact_mixed"):
        r, t = runs[label]
        out[label + "_tight"] = {"seconds": t, "matvecs": r.matvecs, "evals": r.evals.tolist(),
                                 "resid": r.resid.tolist(), "stop_tol_mev": 1e-7}
    ref = np.minimum(runs["exact"][0].evals, runs["exact_mixed"][0].evals)
    out["scoring_evals"] = ref.tolist()
    for label in ("fp64", "mixed"):
        r, t = runs[label]
        err = np.abs(r.evals - ref)
        ok = int((err <= 1e-4).sum())
        out[label] = {"seconds": t, "matvecs": r.matvecs, "evals": r.evals.tolist(), "resid": r.resid.tolist(),
                      "abs_err_vs_scoring": err.tolist(), "within_1e-4": ok, "eigenpairs_per_s": ok / t}
    out["speedup_mixed_vs_fp64"] = out["mixed"]["eigenpairs_per_s"] / max(out["fp64"]["eigenpairs_per_s"], 1e-30)
    return out


def block_leg(s, K, steps, warmup, g):
    """One Hamiltonian on K interleaved FP32 vectors (bs_apply_block, the block
    apply of block Lanczos): the autotuned multi-vector variant, its time per
    launch and H x/s per vector."""
    import torch
    X = torch.randn(s.dim, K, device="cuda", dtype=torch.float32, generator=g)
    Y = s.apply_block(X)  # picks the variant
    for _ in range(warmup):
        s.apply_block(X, Y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        s.apply_block(X, Y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    chosen, tms = s.block_variant_info("f32", K)
    del X, Y
    torch.cuda.empty_cache()
    return {"K": K, "kernel_ms": ms, "hv_per_s_per_vector": K * 1000.0 / ms, "variant": chosen,
            "autotune_ms": {str(v): round(float(t), 3) for v, t in enumerate(tms) if t > 0},
            "what": "0 = one wave per state, 1-4 = one proton row per workgroup, 5-6 = two-row tiles"}


def batch_leg(sps, case, K, steps, warmup, solve, max_restarts, work, n_ops):
    """K Hamiltonians on one basis: interleaved multi-vector matvec (every
    jump read serves K vectors) and, with solve, the lockstep thick-restart
    Lanczos for all K in mixed precision plus Hellmann-Feynman gradients of
    every kept eigenpair, scored against the same solve in FP64."""
    import numpy as np
    import torch
    from bigstickpublick_amd import Batch
    from bigstickpublick_amd.inputs import gaussian_interaction, write_case
    files = []
    for k in range(K):
        c = dataclasses.replace(case)
        if k:
            c.interaction = gaussian_interaction(case.orbits, case.interaction.spe, seed=3000 + k)
            c.name = "%s_h%02d" % (case.name, k)
        write_case(c, work)
        files.append(os.path.join(work, c.name + ".int"))
    t0 = time.time()
    b = Batch(sps, files, case.Z, case.N, case.jz2)
    t_build = time.time() - t0
    D = b.dim
    g = torch.Generator(device="cuda")
    g.manual_seed(99)
    x = torch.randn((D, K), device="cuda", dtype=torch.float32, generator=g)
    y = torch.empty_like(x)
    for _ in range(max(warmup, 1)):
        b.apply(x, y)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    from bigstickpublick_amd.engine import load_library, BS_F32
    import ctypes
    L = load_library()
    for _ in range(steps):
        rc = L.bs_batch_apply(b.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), BS_F32)
        if rc:
            raise RuntimeError(L.bs_last_error().decode())
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    chosen, tms = b.variant_info("f32")
    out = {"K": K, "build_s": t_build, "kernel_ms": ms, "hv_per_s": K / (ms * 1e-3),
           "variant": chosen, "variant_ms": [round(float(v), 3) for v in tms]}
    # explicit-CSR model with K value columns: 4 B index + 4K B values per
    # stored (half) element, read x / read-write y for K vectors
    b_alg = n_ops / 2 * (4 + 4 * K) + 3 * 4 * D * K
    out["b_alg_bytes"] = b_alg
    out["hbm_gbs_alg"] = b_alg / (ms * 1e-3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    if solve:
        thick, niter, nkeep = case.thick, case.niter, case.nkeep
        runs = {}
        for dt, tol in (("f64", 1e-7), ("f32", 1e-5)):
            torch.cuda.synchronize()
            t0 = time.time()
            r = b.lanczos(nkeep=nkeep, niter=niter, thick=thick, max_restarts=max_restarts, dtype=dt, tol=tol,
                          polish=True, want_vectors=(dt == "f32"))
            torch.cuda.synchronize()
            t_solve = time.time() - t0
            t_grad = 0.0
            if dt == "f32":
                t0 = time.time()
                grads = [b.gradient(r.evecs[q]) for q in range(nkeep)]
                torch.cuda.synchronize()
                t_grad = time.time() - t0
                out["gradient_shape"] = [nkeep] + list(grads[0].shape)
            runs[dt] = (r, t_solve, t_grad)
            torch.cuda.empty_cache()
        r64, t64, _ = runs["f64"]
        r32, t32, tg = runs["f32"]
        err = np.abs(r32.evals - r64.evals)
        ok = int((err <= 1e-4).sum())
        out["solve"] = {"fp64_batched": {"seconds": t64, "matvecs": r64.matvecs},
                        "mixed_batched": {"seconds_solve": t32, "seconds_gradients": tg, "matvecs": r32.matvecs,
                                          "within_1e-4": ok, "of": int(err.size),
                                          "max_abs_err_vs_fp64": float(err.max()),
                                          "eigenpairs_per_s": ok / (t32 + tg)},
                        "evals_fp64": r64.evals.tolist()}
    b.close()
    return out


def ensemble_bench(a):
    """The batched ensemble (64 seeded Hamiltonians on the sd 4p4n basis,
    D = 28,503): one interleaved 64-vector H x per step; with --solve, the
    lockstep batched Lanczos (150 steps, lowest 5, FP64 polish) in mixed
    precision plus Hellmann-Feynman gradients of all 5 x 64 eigenpairs, against
    the plain FP64 single-Hamiltonian path solving the same 64 one by one."""
    import numpy as np
    import torch
    from bigstickpublick_amd import Batch, System
    from bigstickpublick_amd.inputs import config_case, write_case
    K = 64
    work = tempfile.mkdtemp(prefix="bsbench_")
    cases = [config_case(2, k) for k in range(K)]
    for c in cases:
        write_case(c, work)
    sps = os.path.join(work, "sd.sps")
    files = [os.path.join(work, c.name + ".int") for c in cases]
    t0 = time.time()
    b = Batch(sps, files, 4, 4, 0)
    t_build = time.time() - t0
    D = b.dim
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.randn((D, K), device="cuda", dtype=torch.float32, generator=g)
    y = torch.empty_like(x)
    import ctypes
    from bigstickpublick_amd.engine import BS_F32, load_library
    L = load_library()
    for _ in range(max(a.warmup, 1)):
        b.apply(x, y)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.time()
    ev0.record()
    for _ in range(a.steps):
        rc = L.bs_batch_apply(b.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), BS_F32)
        if rc:
            raise RuntimeError(L.bs_last_error().decode())
    ev1.record()
    torch.cuda.synchronize()
    wall = time.time() - t0
    ms = ev0.elapsed_time(ev1) / a.steps
    chosen, tms = b.variant_info("f32")
    out = {
        "metric": "H·v/s of the 64-Hamiltonian batched matvec (sd 4p4n ensemble)",
        "value": K * a.steps / wall, "unit": "H·v/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall * 1000.0 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "sd4p4n ensemble of %d seeded Hamiltonians, M=0 D=%d, interleaved FP32 H x" % (K, D),
                   "dim": D, "hamiltonians": K},
        "kernel_ms": ms, "batch_build_s": t_build, "variant": chosen,
        "variant_ms": [round(float(v), 4) for v in tms],
    }
    del x, y
    if a.solve:
        res = {}
        for dt in ("f64", "f32"):
            b.lanczos(nkeep=5, niter=150, dtype=dt, polish=True, want_vectors=(dt == "f32"))  # capture graphs
            torch.cuda.synchronize()
            t0 = time.time()
            r = b.lanczos(nkeep=5, niter=150, dtype=dt, polish=True, want_vectors=(dt == "f32"))
            torch.cuda.synchronize()
            ts = time.time() - t0
            tg = 0.0
            if dt == "f32":
                t0 = time.time()
                grads = [b.gradient(r.evecs[q]) for q in range(5)]
                tg = time.time() - t0
            res[dt] = (r, ts, tg)
        r64, t64b, _ = res["f64"]
        r32, t32, tg = res["f32"]
        err = np.abs(r32.evals - r64.evals)
        ok = int((err <= 1e-4).sum())
        # plain FP64 path: each Hamiltonian on its own (tables built outside the timing)
        systems = [System(sps, f, 4, 4, 0, upload=True) for f in files]
        for s in systems:
            s.autotune("f64")
        torch.cuda.synchronize()
        t0 = time.time()
        single = [s.lanczos(nkeep=5, niter=150, dtype="f64", polish=True).evals for s in systems]
        torch.cuda.synchronize()
        t_single = time.time() - t0
        err_single = np.abs(np.array(single) - r64.evals).max()
        out["solve"] = {
            "plain_fp64_single": {"seconds": t_single, "eigenpairs": 5 * K, "eigenpairs_per_s": 5 * K / t_single,
                                  "max_abs_diff_vs_batched_fp64": float(err_single)},
            "batched_fp64": {"seconds": t64b, "eigenpairs_per_s": 5 * K / t64b},
            "batched_mixed_with_gradients": {"seconds_solve": t32, "seconds_gradients": tg,
                                             "gradient_shape": [5] + list(grads[0].shape),
                                             "within_1e-4": ok, "of": int(err.size),
                                             "max_abs_err_vs_fp64": float(err.max()),
                                             "eigenpairs_per_s": ok / (t32 + tg)},
        }
        out["solve"]["speedup_vs_plain_fp64"] = (out["solve"]["batched_mixed_with_gradients"]["eigenpairs_per_s"]
                                                 / out["solve"]["plain_fp64_single"]["eigenpairs_per_s"])
    b.close()
    print(json.dumps(out))


def main():
    a = parse()
    if a.traffic:
        from bigstickpublick_amd.inputs import config_case
        from bigstickpublick_amd import System
        import re
        case = config_case(a.config)
        work = tempfile.mkdtemp(prefix="bsbench_")
        from bigstickpublick_amd.inputs import write_case
        write_case(case, work)
        s = System(os.path.join(work, case.space + ".sps"), os.path.join(work, case.name + ".int"),
                   case.Z, case.N, case.jz2)
        t = traffic_from_dbs(a.traffic[0], a.traffic[1], s.dim)
        m = re.search(r"k_apply_tiles<float, 512, (\d+), (\d+), (\d+), (true|false)>", t["kernel"])
        t["workload"] = "%s M=0 D=%d FP32" % (case.name, s.dim)
        t["variant"] = [int(m.group(2)), int(m.group(3)), a.variant if a.variant >= 0 else None,
                        int(m.group(4) == "true")]
        if a.variant >= 0:
            t["variant"] = list(System.variants()[a.variant])
        old = {"variants": []}
        if os.path.exists(TRAFFIC_JSON):
            with open(TRAFFIC_JSON) as f:
                old = json.load(f)
            if "variants" not in old:
                old = {"variants": [old]}
        old["variants"] = [e for e in old["variants"] if e.get("variant") != t["variant"]] + [t]
        with open(TRAFFIC_JSON, "w") as f:
            json.dump(old, f, indent=1)
        print(json.dumps({"bytes_per_launch": t["bytes_per_launch"], "kernel": t["kernel"]}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.config == 2:
        ensemble_bench(a)
        return
    import numpy as np
    import torch
    from bigstickpublick_amd import System
    from bigstickpublick_amd.inputs import config_case, write_case

    case = config_case(a.config)
    work = tempfile.mkdtemp(prefix="bsbench_")
    write_case(case, work)
    sps, intf = os.path.join(work, case.space + ".sps"), os.path.join(work, case.name + ".int")
    t0 = time.time()
    s = System(sps, intf, case.Z, case.N, case.jz2)
    t_host = time.time() - t0
    t0 = time.time()
    s.upload()
    torch.cuda.synchronize()
    t_upload = time.time() - t0
    t0 = time.time()
    if a.variant >= 0:
        if not s.set_variant("f32", a.variant):
            raise SystemExit("kernel variant %d does not fit this system" % a.variant)
        variant = a.variant
    else:
        variant = s.autotune("f32")
    t_tune = time.time() - t0
    D = s.dim
    ops = s.op_counts()
    n_ops = ops["ref_total"]
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    x = torch.randn(D, device=dev, dtype=torch.float32, generator=g)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    for _ in range(a.warmup):
        s.apply(x, y)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.time()
    ev0.record(stream)
    for _ in range(a.steps):
        s.apply(x, y)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.time() - t0
    kern_ms = ev0.elapsed_time(ev1) / a.steps
    if a.copy_probe:
        y.copy_(x)
        torch.cuda.synchronize()
    ms_per_step = wall * 1000.0 / a.steps
    b_alg = n_ops / 2 * 8 + 3 * 4 * D
    tr = traffic_from_profiles(System.variants()[variant])
    traffic = tr["bytes_per_launch"] if tr else None
    achieved = traffic / (kern_ms * 1e-3) / 1e9 if traffic else None
    compulsory = 2 * 4 * D
    out = {
        "metric": METRIC,
        "value": a.steps / wall,
        "unit": "H·v/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%s M=0 D=%d FP32 factorized H x (seeded Gaussian TBMEs)" % (case.name, D),
                   "space": case.space, "Z": case.Z, "N": case.N, "dim": D},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic},
        "roofline_basis": "achieved = measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                          "copy-calibrated, %s) / kernel time measured here" % os.path.relpath(TRAFFIC_JSON, ROOT),
        "profiled_kernel_ms": tr["fetch"]["mean_duration_ns"] * 1e-6 if tr else None,
        "compulsory": {"bytes": compulsory, "gbs": compulsory / (kern_ms * 1e-3) / 1e9,
                       "frac": compulsory / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "what": "x read once and y written once (FP32), tables not counted"},
        "stored_matrix_equivalent": {"bytes": b_alg, "gbs": b_alg / (kern_ms * 1e-3) / 1e9,
                                     "what": "explicit Hermitian CSR of the same H: N_ops/2*(4 B value + 4 B index) "
                                             "+ 12*D; not read by this kernel"},
        "kernel_ms": kern_ms,
        "n_ops_ref": n_ops,
        "updates_per_s_engine": (ops["spe"] + ops["pp"] + ops["nn"] + ops["pn"]) / (kern_ms * 1e-3),
        "setup_s": {"host_tables": t_host, "upload": t_upload, "autotune": t_tune},
        "autotune_ms": {"R%d_G%d_xcd%d_pf%d" % v: round(float(t), 3)
                        for v, t in zip(System.variants(), s.autotune_times("f32")) if t > 0},
        "traffic_source": os.path.relpath(TRAFFIC_JSON, ROOT) if traffic else None,
        "kernel_variant": {"rows_per_tile": System.variants()[variant][0],
                           "staged_rows": System.variants()[variant][1],
                           "xcd_map": System.variants()[variant][2],
                           "prefetch": System.variants()[variant][3]},
    }
    if a.block:
        out["block"] = block_leg(s, a.block, a.steps, a.warmup, g)
    if a.solve or a.batch:
        del x, y
        x = y = None
        torch.cuda.empty_cache()
    if a.solve:
        out["solve"] = solve_leg(s, case, a.max_restarts)
        # the FP64 H x of the plain FP64 path (autotuned inside the solve leg)
        xd = torch.randn(D, device=dev, dtype=torch.float64, generator=g)
        yd = torch.empty_like(xd)
        s.apply(xd, yd)
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(3):
            s.apply(xd, yd)
        ev1.record(stream)
        torch.cuda.synchronize()
        out["solve"]["fp64_kernel_ms"] = ev0.elapsed_time(ev1) / 3
        out["solve"]["fp32_kernel_ms"] = kern_ms
        del xd, yd
    if a.batch:
        s.close()
        torch.cuda.empty_cache()
        out["batch"] = batch_leg(sps, case, a.batch, a.steps, a.warmup, a.solve, a.max_restarts, work, n_ops)
        if a.solve and "solve" in out:
            out["batch"]["speedup_vs_fp64_single"] = (out["batch"]["solve"]["mixed_batched"]["eigenpairs_per_s"]
                                                      / out["solve"]["fp64"]["eigenpairs_per_s"])
    if not a.no_cpu:
        try:
            del x, y
            out["cpu_baseline"] = cpu_baseline(sps, intf, case)
        except Exception as e:  # the CPU leg is reported, never fatal
            out["cpu_baseline"] = None
            out["cpu_baseline_error"] = str(e)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
