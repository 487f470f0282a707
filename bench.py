"""Benchmark of the PLR apply/solve hot path on one MI355X.

Workload (N=65536, 10 levels, leaf 64, rank 32, seeded random PLR with
orthonormal factors, sig_j = 10^(-j/4), leaves N(0,1)/8 + 4I): inputs are
resident in HBM before timing.  Three independent matrices are rotated so
that the 256 MB Infinity Cache cannot hold the working set between calls.

Prints ONE JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--m", type=int, default=65536)
    ap.add_argument("--nlvl", type=int, default=10)
    ap.add_argument("--rank", type=int, default=32)
    ap.add_argument("--nmat", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def event_time(torch, fn, reps):
    """Average ms per call of fn() measured with events on the current stream."""
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for i in range(reps):
        fn(i)
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def cpu_baseline(args, log):
    """The reference sources (oracle/_ref, OpenBLAS) on host cores, bounded sample."""
    import numpy as np

    import oracle
    import dplr_amd as D

    ref = oracle.load_ref()
    kind = "reference"
    if ref is None:
        ref = oracle.load_oracle()
        kind = "port"
    lib = ref["_lib"]
    threads = 1
    if kind == "reference":
        try:
            lib_ob = None
            import ctypes
            import glob
            so = glob.glob(os.path.join(os.path.dirname(np.__file__), "..", "scipy.libs",
                                        "libscipy_openblas*.so"))
            if so:
                lib_ob = ctypes.CDLL(so[0])
                lib_ob.scipy_openblas_set_num_threads(1)
        except Exception:
            pass
    A = D.PlrTree.random(args.m, args.nlvl, args.rank, 4.0, 1)
    rng = np.random.default_rng(0)
    B = np.asfortranarray(rng.standard_normal((args.m, args.nrhs)))
    t0 = time.perf_counter()
    reps = 0
    while True:
        D.gepdmm("N", "N", A, B, args.nrhs, api=ref)
        reps += 1
        el = time.perf_counter() - t0
        if el > args.cpu_seconds or reps >= 50:
            break
    rhs_s = reps * args.nrhs / el
    log(f"cpu baseline ({kind}): {reps} x dplrgepdmm {args.nrhs} RHS in {el:.2f}s")
    A.free()
    return {"value": rhs_s, "unit": "RHS/s", "cores": threads, "kind": kind,
            "sample": f"{reps} x dplrgepdmm('N','N') with {args.nrhs} RHS at N={args.m} "
                      f"(reference C sources + OpenBLAS 0.3.28, 1 thread)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import dplr_amd as D

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    torch.cuda.init()
    m, nrhs = args.m, args.nrhs
    log(f"generating {args.nmat} PLR matrices N={m} L={args.nlvl} r={args.rank}")
    plans = []
    for k in range(args.nmat):
        t = D.PlrTree.random(m, args.nlvl, args.rank, 4.0, 1 + k)
        plans.append(D.Plan(t))
        t.free()
    abytes = plans[0].nbytes
    rng = np.random.default_rng(0)
    X = torch.tensor(rng.standard_normal((nrhs, m)), device=dev)
    Y = torch.zeros_like(X)
    x1 = X[0].contiguous()
    y1 = torch.zeros(m, dtype=torch.float64, device=dev)

    # --- single-RHS matvec (HBM-bound) ------------------------------------
    mv = lambda i: plans[i % len(plans)].gepmv("N", 1.0, x1, 0.0, y1)
    for i in range(args.warmup * 3):
        mv(i)
    reps = max(args.steps * 5, 30)
    ms_mv = event_time(torch, mv, reps)
    mv_bytes = abytes + 2 * 8 * m
    mv_gbs = mv_bytes / (ms_mv * 1e-3) / 1e9
    log(f"matvec 1 RHS: {ms_mv * 1e3:.1f} us  {mv_gbs:.0f} GB/s")

    # --- 64-RHS FP64 apply (timed step) ------------------------------------
    step = lambda i: plans[i % len(plans)].gepdmm("N", 1.0, X, 0.0, Y)
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_step = event_time(torch, step, args.steps)
    wall = time.perf_counter() - t0
    rhs_s = nrhs / (ms_step * 1e-3)
    flops = nrhs * (2.0 * plans[0].leaf * m + 4.0 * m * sum(plans[0].ranks))
    log(f"apply {nrhs} RHS: {ms_step * 1e3:.1f} us/step  {rhs_s:.0f} RHS/s  "
        f"{flops / (ms_step * 1e-3) / 1e12:.2f} TFLOP/s")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args, log)
        except Exception as exc:  # the baseline is informative only
            log(f"cpu baseline failed: {exc}")

    out = {
        "metric": "PLR apply+solve RHS/s: mixed-prec vs FP64 plain at rel err <=1e-10, N=64K",
        "value": rhs_s,
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"FP64 PLR apply, {nrhs} RHS, N={m}, L={args.nlvl}, leaf "
                               f"{m >> args.nlvl}, rank {args.rank}",
                   "global_batch": nrhs, "matrices_rotated": args.nmat},
        "roofline": {"bound": "hbm", "achieved": mv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": mv_gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "single-RHS matvec (mv_reduce + mv_expand)",
                     "bytes_per_launch": mv_bytes, "us_per_launch": ms_mv * 1e3},
        "cpu_baseline": cpu,
        "wall_s_timed": wall,
    }
    print(json.dumps(out), flush=True)
    for p in plans:
        p.close()


if __name__ == "__main__":
    main()
