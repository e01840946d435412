"""Benchmark: factorized M-scheme H x on one MI355X.

Workload: pf valence space (0f7/2 1p3/2 0f5/2 1p1/2), 6 protons + 6 neutrons,
M = 0, D = 109,954,620, seeded Gaussian two-body ensemble; FP32 vectors and
FP32 matrix elements.  One step = one full H x over the resident vector.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--solve]

Prints one JSON line.  value = H x per second on the whole job (one GPU).
roofline.achieved = B_alg / (mean kernel time), B_alg = N_ops/2 * 8 + 3*4*D
bytes: the traffic of an explicit Hermitian CSR of the same H (FP32 value +
int32 index per stored nonzero, read x / read-write y), with N_ops the
reference's own operation count (src/bigstick_main.f90:1150-1192), which this
engine reproduces.
"""
import argparse
import json
import os
import re
import subprocess
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
    ap.add_argument("--cpu-iters", type=int, default=2, help="reference Lanczos steps in the CPU sample")
    ap.add_argument("--config", type=int, default=4, help="benchmark configuration (4 = pf 6p6n)")
    return ap.parse_args()


def traffic_from_profiles():
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), scaled by a copy-kernel calibration; committed under profiles/."""
    p = os.path.join(ROOT, "profiles", "apply_rows_pf6p6n_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("bytes_per_launch")


def cpu_baseline(case, iters):
    """The reference executable (oracle/_ref, built from the reference's
    sources) on this host's cores: a Lanczos run of `iters` steps on the same
    workload; the matvec rate comes from its own end-of-run report
    ('Time total in H mat-vec multiply', src/bigstick_main.f90)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "bigstick.x")
    if not os.path.exists(exe):
        return None
    from bigstickpublick_amd.inputs import menu_script, write_case
    ncores = len(os.sched_getaffinity(0))
    threads = min(ncores, 16)
    with tempfile.TemporaryDirectory() as d:
        write_case(case, d)
        c = case
        c.niter = iters
        c.nkeep = 1
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t0 = time.time()
        with open(os.path.join(d, "out.log"), "w") as fo:
            r = subprocess.run([exe], input=menu_script(c, "n", "lf"), text=True, stdout=fo,
                               stderr=subprocess.STDOUT, cwd=d, env=env, timeout=3000)
        wall = time.time() - t0
        txt = open(os.path.join(d, "out.log")).read()
    m = re.search(r"Time total in H mat-vec multiply\s*:\s*([0-9.Ee+-]+)", txt)
    if r.returncode != 0 or not m:
        return None
    tmv = float(m.group(1))
    return {"value": iters / tmv, "unit": "H·v/s", "cores": threads, "kind": "reference",
            "sample": "reference BIGSTICK (oracle/_ref, OpenMP) on the same pf 6p6n inputs: "
                      "%d Lanczos iterations = %d H x, %.1f s in H x of %.1f s wall (setup included "
                      "in wall only)" % (iters, iters, tmv, wall)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
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
    ms_per_step = wall * 1000.0 / a.steps
    b_alg = n_ops / 2 * 8 + 3 * 4 * D
    achieved = b_alg / (kern_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles()
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "hbm_gbs_alg": achieved,
        "kernel_ms": kern_ms,
        "n_ops_ref": n_ops,
        "updates_per_s_engine": (ops["spe"] + ops["pp"] + ops["nn"] + ops["pn"]) / (kern_ms * 1e-3),
        "b_alg_bytes": b_alg,
        "setup_s": {"host_tables": t_host, "upload": t_upload, "autotune": t_tune},
        "kernel_variant": {"rows_per_tile": System.variants()[variant][0],
                           "staged_rows": System.variants()[variant][1]},
    }
    if not a.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(case, a.cpu_iters)
        except Exception as e:  # the CPU leg is reported, never fatal
            out["cpu_baseline"] = None
            out["cpu_baseline_error"] = str(e)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
