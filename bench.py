"""Benchmark of the PLR apply/solve hot path on one MI355X.

Workload: N=65536, 10 levels, leaf 64, rank 32, 64 right-hand sides; seeded
random PLR matrices (orthonormal U, V from Householder QR of N(0,1)
matrices, sig_j = 10^(-j/4), leaves N(0,1)/8 + 4I).  Inputs are resident in
HBM before timing; three independent matrices are rotated so the 256 MB
Infinity Cache cannot hold the working set between calls.

One step = apply and solve for the 64 right-hand sides, both to relative
error <= 1e-10:  Y = A X (FP64 apply) and Z = inv(A) B.
  * mixed  (the reported value): one call (dplrx_apply_solve_refined_tol):
    Z from f16-pair factors with FP32 accumulation on MFMA, refined with FP64
    residuals until its estimated relative error is <= 1e-10, the apply
    computed in the same pass over the factors as the first residual; the
    same step as two calls (FP64 apply, then dplrx_solve_refined_tol) is
    reported beside it;
  * plain FP64: the FP64 apply and the FP64 factored solve.
The refined solve stops on its own error estimate; Y and Z are checked
against the separate FP64 apply and the FP64 solve before timing, and
`value` is withheld (null, exit status 1) when an error exceeds 1e-10.
Factorizations are done once per matrix outside the step and timed
separately ("factor_ms"); "incl_factor" gives the rates when every step also
factors its matrix.

Prints ONE JSON line.
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
TARGET_ERR = 1e-10


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
    ap.add_argument("--lp", default="f16x2", choices=["f16x2", "f16", "bf16"])
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="runs per thread count of the CPU baseline (median reported)")
    ap.add_argument("--maxit", type=int, default=8, help="most refined solves per call")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB -> 1 GiB device copy (torch.neg) five times, the "
                         "known-bytes kernel the HBM counters are scaled by, and exit")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-count", type=int, default=512,
                    help="matrices in the batched leg (0 skips it)")
    ap.add_argument("--precision-legs", default="65536:32,131072:64",
                    help="N:rank pairs of the mixed-precision comparison ('' skips it)")
    ap.add_argument("--grad-m", type=int, default=16384,
                    help="order of the Gaussian-kernel gradient leg (0 skips it)")
    return ap.parse_args()


def measured_traffic(name="matvec_fp64_N65536_L10_r32.md", key="traffic_bytes_per_launch"):
    """HBM bytes per launch from a committed counter summary (FETCH_SIZE and
    WRITE_SIZE runs scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def lp_solve_bytes(m, leaf, ranks, nrhs, lp, first):
    """Algorithmic HBM bytes of one low-precision solve of nrhs columns: V, Ut
    and the leaf inverses in 16 bit (two planes for f16x2), the capacitance
    inverses (FP64 on the top two levels, FP32 below), sig in FP32, the
    right-hand side read once in FP64 and x written (first solve) or read and
    written (later ones) in FP64.  Ranks padded to multiples of 32."""
    Rp = [((r + 31) // 32) * 32 for r in ranks]
    planes = 2 if lp == "f16x2" else 1
    fac = 2 * planes * (2 * m * sum(Rp) + m * leaf)
    kinv = sum((1 << d) * (2 * R) ** 2 * (8 if d < 2 else 4) for d, R in enumerate(Rp))
    sig = 4 * sum((2 << d) * R for d, R in enumerate(Rp))
    io = 8 * m * nrhs * (2 if first else 3)
    return fac + kinv + sig + io


def event_ms(torch, fn, reps):
    """Average ms per call of fn(i), HIP events on the current stream (the
    stream every library call in this script is launched on)."""
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for i in range(reps):
        fn(i)
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(args, log):
    """The reference sources (oracle/_ref, linked to the OpenBLAS inside SciPy)
    on the host: apply + explicit inverse + inverse applied for one 64-RHS
    batch on one matrix of the workload shape (a bounded sample), timed
    with a monotonic clock, the median of `cpu_reps` runs, once with one
    BLAS thread and once with every host thread this process may use (the
    reference's recursion is serial; only its BLAS calls thread)."""
    import ctypes
    import glob

    import numpy as np

    import dplr_amd as D
    import oracle

    api = oracle.load_ref()
    kind, blas, set_threads = "reference", None, None
    if api is None:
        api = oracle.load_oracle()
        kind = "port"
        blas = "none (the oracle's own dense loops)"
    else:
        so = glob.glob(os.path.join(os.path.dirname(np.__file__), "..", "scipy.libs",
                                    "libscipy_openblas*.so"))
        if so:
            ob = ctypes.CDLL(so[0])
            cfg = ob.scipy_openblas_get_config
            cfg.restype = ctypes.c_char_p
            blas = cfg().decode()
            set_threads = ob.scipy_openblas_set_num_threads
    all_threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    A = D.PlrTree.random(args.m, args.nlvl, args.rank, 4.0, 1)
    rng = np.random.default_rng(0)
    X = np.asfortranarray(rng.standard_normal((args.m, args.nrhs)))

    def once():
        t0 = time.monotonic()
        D.gepdmm("N", "N", A, X, args.nrhs, api=api)
        t1 = time.monotonic()
        B, info = D.geinv(A, 0.0, api=api)
        t2 = time.monotonic()
        D.gepdmm("N", "N", B, X, args.nrhs, api=api)
        t3 = time.monotonic()
        B.free()
        return t1 - t0, t2 - t1, t3 - t2

    legs = {}
    for nth in ([1, all_threads] if set_threads and all_threads > 1 else [1]):
        if set_threads:
            set_threads(nth)
        runs = [once() for _ in range(args.cpu_reps)]
        tot = sorted(sum(r) for r in runs)
        med = tot[len(tot) // 2]
        legs[nth] = {"threads": nth, "runs": len(runs), "median_s": med,
                     "min_s": tot[0], "max_s": tot[-1], "rhs_s": args.nrhs / med,
                     "apply_median_s": sorted(r[0] for r in runs)[len(runs) // 2],
                     "inverse_median_s": sorted(r[1] for r in runs)[len(runs) // 2],
                     "apply_inverse_median_s": sorted(r[2] for r in runs)[len(runs) // 2]}
        log(f"cpu ({kind}, {nth} thread(s)): median {med:.2f} s over {len(runs)} runs "
            f"-> {args.nrhs / med:.1f} RHS/s")
    if set_threads:
        set_threads(1)
    A.free()
    best = max(legs.values(), key=lambda v: v["rhs_s"])
    return {"value": best["rhs_s"], "unit": "RHS/s", "cores": best["threads"], "kind": kind,
            "sample": f"one {args.nrhs}-RHS batch on one N={args.m} matrix: dplrgepdmm(A) + "
                      f"dplrgeinv + dplrgepdmm(inv(A)), median of {args.cpu_reps} runs",
            "legs": list(legs.values()), "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "blas": blas}


def batch_bytes(plans, nrhs, es, ev):
    """Algorithmic HBM bytes of one batched apply + solve: per matrix U, V,
    the leaves and sig for the apply, V, Ut, the leaf inverses, sig and the
    capacitance inverses for the solve (factor entries of es bytes, sig and
    inv(K) of ev bytes), and the right-hand sides read and results written
    once per operation (ev bytes)."""
    tot = 0
    for P in plans:
        m, b, R = P.m, P.leaf, P.ranks
        uv = m * sum(R) * es
        sig = sum((2 << d) * r for d, r in enumerate(R)) * ev
        kinv = sum((1 << d) * (2 * r) ** 2 for d, r in enumerate(R)) * ev
        io = 2 * m * nrhs * ev
        tot += (2 * uv + m * b * es + sig + io) + (2 * uv + m * b * es + sig + kinv + io)
    return tot


def batch_specs(count, seed=7):
    """Shapes of the batched workload: N log-uniform in [1024, 16384] snapped
    to leaf * 2^L, leaf from {64, 128, 256}, rank uniform in {8..48} (at most
    the order of the smallest off-diagonal block)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        leaf = int(rng.choice([64, 128, 256]))
        n = float(np.exp(rng.uniform(np.log(1024), np.log(16384))))
        L = max(1, int(round(np.log2(n / leaf))))
        while (leaf << L) > 16384:
            L -= 1
        while (leaf << L) < 1024:
            L += 1
        r = min(int(rng.integers(8, 49)), leaf // 2)
        specs.append((leaf << L, L, r))
    return specs


def batch_leg(args, torch, D, log):
    """Many independent matrices in one go: apply + factored solve of 16 RHS
    per matrix for every matrix of the batch, one launch per phase, replayed
    from a HIP graph, in FP64, in FP32 (FP32 factors on FP32 MFMA), and with
    BF16 factors refined by FP32 residuals (1 to 3 solves); tiles from the
    autotuner.  Per-matrix latency from the same work done one matrix at a
    time."""
    import numpy as np
    nrhs = 16
    specs = batch_specs(args.batch_count)
    t0 = time.perf_counter()
    plans = []
    for k, (m, L, r) in enumerate(specs):
        t = D.PlrTree.random(m, L, r, 4.0, 1000 + k)
        plans.append(D.Plan(t))
        t.free()
    for P in plans:
        assert P.factor() == 0
    torch.cuda.synchronize()
    log(f"batch: {len(plans)} matrices generated, uploaded and factored in "
        f"{time.perf_counter() - t0:.1f} s")
    B = D.Batch(plans)
    g = torch.Generator(device="cpu").manual_seed(11)
    X = [torch.randn(nrhs, m, generator=g, dtype=torch.float64).to("cuda") for (m, _, _) in specs]
    Y = [torch.zeros_like(x) for x in X]
    Z = [torch.zeros_like(x) for x in X]

    # autotuner: the tile shape (rows per reduction task) for this batch
    tile, tile_ms = B.tune(X, Y, reps=3)
    log(f"batch: tuner picked {tile}-row tiles ("
        + ", ".join(f"{k}: {v:.2f} ms" for k, v in tile_ms.items()) + ")")

    def both(_):
        B.gepdmm(X, Y)
        B.solve(X, Z)
    for i in range(3):
        both(i)
    ms = event_ms(torch, both, 10)
    ms_apply = event_ms(torch, lambda i: B.gepdmm(X, Y), 10)
    ms_solve = event_ms(torch, lambda i: B.solve(X, Z), 10)
    # accuracy of every solve through the single-plan FP64 apply
    err = 0.0
    for P, x, z in zip(plans, X, Z):
        back = torch.zeros_like(x)
        P.gepdmm("N", 1.0, z, 0.0, back)
        err = max(err, float((back - x).norm() / x.norm()))
    # per-matrix latency of the same work, one matrix at a time
    lat = []
    for P, x, y, z in zip(plans, X, Y, Z):
        def one(_):
            P.gepdmm("N", 1.0, x, 0.0, y)
            P.solve("N", x, z)
        for i in range(2):
            one(i)
        lat.append(event_ms(torch, one, 3))
    lat = np.array(lat)
    total = len(plans) * nrhs
    by64 = batch_bytes(plans, nrhs, 8, 8)
    out = {"matrices": len(plans), "rhs_per_matrix": nrhs,
           "rhs_s": total / (ms * 1e-3), "ms_per_batch": ms,
           "roofline": {"bound": "hbm", "bytes_per_batch": by64,
                        "achieved": by64 / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": by64 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "apply_ms": ms_apply, "solve_ms": ms_solve,
           "max_rel_residual": err, "dtype": "f64",
           "tile_rows": tile, "tuner_ms": tile_ms,
           "per_matrix_ms_p50": float(np.percentile(lat, 50)),
           "per_matrix_ms_p99": float(np.percentile(lat, 99)),
           "one_at_a_time_rhs_s": total / (lat.sum() * 1e-3),
           "sizes": f"N in [{min(s[0] for s in specs)}, {max(s[0] for s in specs)}], "
                    f"rank in [{min(s[2] for s in specs)}, {max(s[2] for s in specs)}], "
                    "leaf in {64, 128, 256}"}
    log(f"batch: {out['rhs_s']:.0f} RHS/s ({ms:.2f} ms per batch of {total} RHS; apply "
        f"{ms_apply:.2f} ms, solve {ms_solve:.2f} ms; {by64 / 1e9:.2f} GB algorithmic, "
        f"{out['roofline']['frac']:.1%} of HBM); max residual {err:.1e}; one matrix at a "
        f"time p50 {out['per_matrix_ms_p50']:.3f} ms p99 {out['per_matrix_ms_p99']:.3f} ms")

    # FP32 (FP32 factors, FP32 MFMA) and BF16 factors refined with FP32
    # residuals, on the same matrices and right-hand sides
    for P in plans:
        assert P.factor_single("f32") == 0 and P.factor_single("bf16") == 0
    X32 = [x.float() for x in X]
    Y32 = [torch.zeros_like(x) for x in X32]
    Z32 = [torch.zeros_like(x) for x in X32]
    tile32, tile32_ms = B.tune(X32, Y32, reps=3, prec="f32")

    def both32(_):
        B.gepdmm(X32, Y32, prec="f32")
        B.solve(X32, Z32, prec="f32")
    for i in range(3):
        both32(i)
    ms32 = event_ms(torch, both32, 10)
    ms32_apply = event_ms(torch, lambda i: B.gepdmm(X32, Y32, prec="f32"), 10)
    ms32_solve = event_ms(torch, lambda i: B.solve(X32, Z32, prec="f32"), 10)

    def rel_vs64(Zs):
        return max(float((z.double() - z64).norm() / z64.norm()) for z, z64 in zip(Zs, Z))
    err32 = rel_vs64(Z32)
    lat32 = []
    for P, x, y, z in zip(plans, X32, Y32, Z32):
        def one32(_):
            P.gepdmm_single("N", 1.0, x, 0.0, y)
            P.solve_single("N", x, z)
        for i in range(2):
            one32(i)
        lat32.append(event_ms(torch, one32, 3))
    lat32 = np.array(lat32)
    by32 = batch_bytes(plans, nrhs, 4, 4)
    f32 = {"dtype": "f32", "rhs_s": total / (ms32 * 1e-3), "ms_per_batch": ms32,
           "roofline": {"bound": "hbm", "bytes_per_batch": by32,
                        "achieved": by32 / (ms32 * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": by32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "apply_ms": ms32_apply, "solve_ms": ms32_solve, "rel_err_vs_fp64_solve": err32,
           "tile_rows": tile32, "tuner_ms": tile32_ms,
           "per_matrix_ms_p50": float(np.percentile(lat32, 50)),
           "per_matrix_ms_p99": float(np.percentile(lat32, 99))}
    log(f"batch FP32: {f32['rhs_s']:.0f} RHS/s ({ms32:.2f} ms; apply {ms32_apply:.2f} ms, solve "
        f"{ms32_solve:.2f} ms, tile {tile32}; {f32['roofline']['frac']:.1%} of HBM); error vs "
        f"FP64 {err32:.1e}; one at a time p50 {f32['per_matrix_ms_p50']:.3f} ms p99 "
        f"{f32['per_matrix_ms_p99']:.3f} ms")
    refined = []
    for iters in (1, 2, 3):
        ZR = [torch.zeros_like(x) for x in X32]

        def both_r(_):
            B.gepdmm(X32, Y32, prec="f32")
            B.solve_refined(X32, ZR, iters)
        for i in range(3):
            both_r(i)
        msr = event_ms(torch, both_r, 10)
        msr_solve = event_ms(torch, lambda i: B.solve_refined(X32, ZR, iters), 10)
        errr = rel_vs64(ZR)
        refined.append({"dtype": "bf16 refined with f32 residuals", "solves": iters,
                        "rhs_s": total / (msr * 1e-3), "ms_per_batch": msr,
                        "solve_ms": msr_solve, "rel_err_vs_fp64_solve": errr})
        log(f"batch BF16 refined ({iters} solves): {total / (msr * 1e-3):.0f} RHS/s ({msr:.2f} ms; "
            f"solve {msr_solve:.2f} ms); error vs FP64 {errr:.1e}")
    out["f32"] = f32
    out["bf16_refined"] = refined
    B.close()
    for P in plans:
        P.close()
    return out


def precision_leg(args, torch, D, log):
    """Mixed precision against the plain FP64 solve, per (N, rank) and per
    16-bit type: refined solves needed for rel err <= 1e-10 against the FP64
    solve, the error reached, and the solve times (64 RHS)."""
    import numpy as np
    out = []
    nrhs = args.nrhs
    for item in [x for x in args.precision_legs.split(",") if x]:
        m, r = (int(v) for v in item.split(":"))
        nlvl = int(round(np.log2(m // 64)))
        t = D.PlrTree.random(m, nlvl, r, 4.0, 7)
        P = D.Plan(t)
        t.free()
        assert P.factor() == 0
        g = torch.Generator(device="cpu").manual_seed(3)
        B = torch.randn(nrhs, m, generator=g, dtype=torch.float64).to("cuda")
        Z64 = torch.zeros_like(B)
        Z = torch.zeros_like(B)
        for _ in range(3):  # eager, captured, replayed
            P.solve("N", B, Z64)
        ms64 = event_ms(torch, lambda i: P.solve("N", B, Z64), 5)
        for lp in ("f16x2", "f16", "bf16"):
            if lp == "f16x2" and max(P.ranks) > 32:
                continue  # pair storage covers ranks up to 32 at these orders
            assert P.factor_mixed(lp) == 0
            iters, est = P.solve_refined_tol(B, Z, tol=TARGET_ERR, maxit=args.maxit)
            torch.cuda.synchronize()
            err = float((Z - Z64).norm() / Z64.norm())
            ms = event_ms(torch, lambda i: P.solve_refined_tol(B, Z, tol=TARGET_ERR,
                                                               maxit=args.maxit), 5)
            ms1 = event_ms(torch, lambda i: P.solve_refined(B, Z, 1), 5)
            rec = {"m": m, "rank": r, "lowprec": lp, "refine_solves": iters,
                   "error_estimate": est, "rel_err": err,
                   "target_met": err <= TARGET_ERR and est <= TARGET_ERR,
                   "mixed_solve_ms": ms, "single_lowprec_solve_ms": ms1, "fp64_solve_ms": ms64,
                   "speedup_vs_fp64_solve": ms64 / ms}
            out.append(rec)
            log(f"precision N={m} r={r} {lp}: {iters} solves, estimate {est:.1e}, error "
                f"{err:.1e}, {ms:.3f} ms vs FP64 {ms64:.3f} ms ({ms64 / ms:.2f}x); one {lp} "
                f"solve {ms1:.3f} ms")
        P.close()
        torch.cuda.empty_cache()
    return out


def grad_leg(args, torch, D, log):
    """Gaussian-kernel PLR (x ~ U[0,1] sorted, l = 0.05, nugget 1e-2,
    compressed at tol 1e-10 with leaf 128): forward y = A^{-1} b and
    q = b^T y, backward dq/db and dq/d(U, V, sig, leaves) by the adjoint
    kernels (transposed solve + projected outer product), FP64."""
    import ctypes
    import numpy as np
    from dplr_amd.autograd import plr_solve
    m, leaf = args.grad_m, 128
    nlvl = int(round(np.log2(m // leaf)))
    x = np.sort(np.random.default_rng(0).uniform(0, 1, m))
    A = np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * 0.05 ** 2))
    A[np.diag_indices(m)] += 1e-2
    A = np.asfortranarray(A)
    T = D.PlrTree(m)
    info = ctypes.c_int(0)
    t0 = time.perf_counter()
    D.lib.dplrd2gep(m, A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m, T.ref, 1e-10, nlvl,
                    ctypes.byref(info))
    t_cmp = time.perf_counter() - t0
    del A
    assert info.value == 0
    P = D.Plan(T)
    ranks = P.ranks
    T.free()
    t0 = time.perf_counter()
    assert P.factor() == 0
    torch.cuda.synchronize()
    t_fac = time.perf_counter() - t0
    b = torch.tensor(np.random.default_rng(2).standard_normal((1, m)), device="cuda")

    def fwd(_):
        return (b * plr_solve(b, P)).sum()

    def fwd_bwd(_):
        bb = b.clone().requires_grad_(True)
        q = (bb * plr_solve(bb, P)).sum()
        q.backward()
        return bb.grad

    for i in range(3):
        fwd_bwd(i)
    ms_f = event_ms(torch, fwd, 10)
    ms_fb = event_ms(torch, fwd_bwd, 10)
    # consistency: dq/db = y + A^{-T} b and, for the symmetric kernel, = 2 y
    g = fwd_bwd(0)
    y = torch.zeros_like(b)
    P.solve("N", b, y)
    err = float((g - 2 * y).norm() / (2 * y).norm())
    out = {"m": m, "leaf": leaf, "levels": nlvl, "ranks": ranks, "dtype": "f64",
           "compress_s": t_cmp, "factor_s": t_fac, "forward_ms": ms_f,
           "forward_backward_ms": ms_fb, "grad_b_vs_2y": err}
    log(f"gradient leg: N={m} ranks {ranks}; compress {t_cmp:.1f} s, factor {t_fac * 1e3:.1f} ms; "
        f"forward {ms_f:.3f} ms, forward+backward {ms_fb:.3f} ms; |dq/db - 2y|/|2y| = {err:.1e}")

    # FP32: the same forward and backward on the plan's FP32 factors
    assert P.factor_single("f32") == 0
    b32 = b.float()
    q64 = float(fwd(0))
    g64 = g.clone()
    fac64 = [x.clone() for x in P.last_factor_grad]

    def fwd32(_):
        return (b32 * plr_solve(b32, P)).sum()

    def fwd_bwd32(_):
        bb = b32.clone().requires_grad_(True)
        q = (bb * plr_solve(bb, P)).sum()
        q.backward()
        return bb.grad
    for i in range(3):
        fwd_bwd32(i)
    ms_f32 = event_ms(torch, fwd32, 10)
    ms_fb32 = event_ms(torch, fwd_bwd32, 10)
    g32 = fwd_bwd32(0)
    fac32 = P.last_factor_grad
    q32 = float(fwd32(0))
    fac_err = max(float((a.double() - b_).norm() / b_.norm()) for a, b_ in zip(fac32, fac64))
    out["f32"] = {"forward_ms": ms_f32, "forward_backward_ms": ms_fb32,
                  "q_rel_err_vs_fp64": abs(q32 - q64) / abs(q64),
                  "grad_b_rel_err_vs_fp64": float((g32.double() - g64).norm() / g64.norm()),
                  "factor_grad_rel_err_vs_fp64": fac_err}
    log(f"gradient leg FP32: forward {ms_f32:.3f} ms, forward+backward {ms_fb32:.3f} ms; vs FP64: "
        f"q {out['f32']['q_rel_err_vs_fp64']:.1e}, dq/db {out['f32']['grad_b_rel_err_vs_fp64']:.1e}, "
        f"factor gradients {fac_err:.1e}")
    P.close()
    return out


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
    if args.copy_calibration:
        # 1 GiB read and 1 GiB written per launch: the scale of FETCH_SIZE
        # and WRITE_SIZE for this kernel gives the counters' units
        x = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        for _ in range(5):
            torch.neg(x, out=y)
        ms = event_ms(torch, lambda i: torch.neg(x, out=y), 5)
        print(json.dumps({"copy_calibration": True, "bytes_read": 1 << 30,
                          "bytes_written": 1 << 30, "ms": ms,
                          "GB_s": 2 * (1 << 30) / (ms * 1e-3) / 1e9}), flush=True)
        return
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
    B = torch.tensor(rng.standard_normal((nrhs, m)), device=dev)
    Y = torch.zeros_like(X)
    Z = torch.zeros_like(X)
    Z64 = [torch.zeros_like(X) for _ in plans]
    x1 = X[0].contiguous()
    y1 = torch.zeros(m, dtype=torch.float64, device=dev)

    # --- factorizations (outside the step) --------------------------------
    t0 = time.perf_counter()
    for P in plans:
        assert P.factor() == 0
    torch.cuda.synchronize()
    f64_ms = (time.perf_counter() - t0) * 1e3 / len(plans)
    t0 = time.perf_counter()
    for P in plans:
        assert P.factor_mixed(args.lp) == 0
    torch.cuda.synchronize()
    lp_ms = (time.perf_counter() - t0) * 1e3 / len(plans)
    # again, with every buffer in place (the cost of refactoring a matrix)
    warm = {}
    for kind in ("fp64", args.lp):
        t0 = time.perf_counter()
        for rep in range(2):
            for P in plans:
                assert (P.factor() if kind == "fp64" else P.factor_mixed(args.lp)) == 0
        torch.cuda.synchronize()
        warm[kind] = (time.perf_counter() - t0) * 1e3 / (2 * len(plans))
    log(f"factor: fp64 {f64_ms:.2f} ms, {args.lp} {lp_ms:.2f} ms first (host-timed, per matrix); "
        f"again: fp64 {warm['fp64']:.2f} ms, {args.lp} {warm[args.lp]:.2f} ms")

    # --- single-RHS matvec (HBM-bound kernel pair) -------------------------
    mv = lambda i: plans[i % len(plans)].gepmv("N", 1.0, x1, 0.0, y1)
    for i in range(6):
        mv(i)
    # median of three runs of 60 calls (the first run after the factorizations
    # pays for the clocks coming back up)
    ms_mv = float(np.median([event_ms(torch, mv, 60) for _ in range(3)]))
    mv_bytes = abytes + 2 * 8 * m
    mv_gbs = mv_bytes / (ms_mv * 1e-3) / 1e9
    log(f"matvec 1 RHS: {ms_mv * 1e3:.1f} us  {mv_gbs:.0f} GB/s ({mv_gbs / HBM_PEAK_GBS:.1%} of HBM)")

    # --- legs of the step ---------------------------------------------------
    apply = lambda i: plans[i % len(plans)].gepdmm("N", 1.0, X, 0.0, Y)
    solve64 = lambda i: plans[i % len(plans)].solve("N", B, Z)
    # every plan's first call runs eagerly and the second is captured into a
    # HIP graph (plr_run_cached); warm up past both for each plan
    for i in range(3 * len(plans)):
        apply(i)
        solve64(i)
    ms_apply = event_ms(torch, apply, 10)
    ms_solve64 = event_ms(torch, solve64, 10)
    flops = nrhs * (2.0 * plans[0].leaf * m + 4.0 * m * sum(plans[0].ranks))
    log(f"apply {nrhs} RHS: {ms_apply * 1e3:.0f} us ({flops / (ms_apply * 1e-3) / 1e12:.1f} TFLOP/s); "
        f"fp64 solve: {ms_solve64 * 1e3:.0f} us")

    # the refined solve stops on its own error estimate; its result is also
    # checked here against each matrix's FP64 solve
    for k, P in enumerate(plans):
        P.solve("N", B, Z64[k])
    torch.cuda.synchronize()
    iters, ests, errs = [], [], []
    for k, P in enumerate(plans):
        its, est = P.solve_refined_tol(B, Z, tol=TARGET_ERR, maxit=args.maxit)
        torch.cuda.synchronize()
        iters.append(its)
        ests.append(est)
        errs.append(((Z - Z64[k]).norm() / Z64[k].norm()).item())
    err, est = max(errs), max(ests)
    target_met = err <= TARGET_ERR and est <= TARGET_ERR
    log(f"mixed ({args.lp}): {max(iters)} refined solves, estimate {est:.2e}, rel err vs the "
        f"FP64 solve {err:.2e} ({'met' if target_met else 'NOT met'})")
    mixed = lambda i: plans[i % len(plans)].solve_refined_tol(B, Z, tol=TARGET_ERR,
                                                              maxit=args.maxit)
    for i in range(2 * len(plans)):
        mixed(i)
    ms_mixed = event_ms(torch, mixed, 10)
    # the same number of solves without the estimate's read-backs
    nfix = max(iters)
    fixed = lambda i: plans[i % len(plans)].solve_refined(B, Z, nfix)
    for i in range(len(plans)):
        fixed(i)
    ms_fixed = event_ms(torch, fixed, 10)
    # one low-precision solve (its column maxima of B included) on the HBM
    # roofline: the refinement's dominant kernels
    one = lambda i: plans[i % len(plans)].solve_refined(B, Z, 1)
    for i in range(len(plans)):
        one(i)
    ms_one = event_ms(torch, one, 20)
    lp_bytes = lp_solve_bytes(m, plans[0].leaf, plans[0].ranks, nrhs, args.lp, True)
    lp_gbs = lp_bytes / (ms_one * 1e-3) / 1e9
    log(f"one {args.lp} solve: {ms_one * 1e3:.0f} us, {lp_bytes / 1e6:.1f} MB algorithmic -> "
        f"{lp_gbs:.0f} GB/s ({lp_gbs / HBM_PEAK_GBS:.1%} of HBM)")

    # --- the timed step ------------------------------------------------------
    # one call per step: Y = A X (FP64) and Z = inv(A) B refined to 1e-10,
    # the apply riding on the first residual product
    # (dplrx_apply_solve_refined_tol); checked first against the separate
    # FP64 apply and the FP64 solve
    Ya = torch.zeros_like(X)
    step_err = 0.0
    for k, P in enumerate(plans):
        P.apply_solve_refined_tol(X, Ya, B, Z, tol=TARGET_ERR, maxit=args.maxit)
        P.gepdmm("N", 1.0, X, 0.0, Y)
        torch.cuda.synchronize()
        step_err = max(step_err, ((Ya - Y).norm() / Y.norm()).item(),
                       ((Z - Z64[k]).norm() / Z64[k].norm()).item())
    target_met = target_met and step_err <= TARGET_ERR
    log(f"apply+solve in one call: largest rel err vs the separate FP64 apply / solve {step_err:.2e}")

    def step(i):
        plans[i % len(plans)].apply_solve_refined_tol(X, Ya, B, Z, tol=TARGET_ERR,
                                                      maxit=args.maxit)

    def step_apart(i):
        apply(i)
        mixed(i)

    for i in range(2 * len(plans)):
        step_apart(i)
    ms_step_apart = event_ms(torch, step_apart, args.steps)
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_step = event_ms(torch, step, args.steps)
    wall = time.perf_counter() - t0
    # per-step distribution (each step bracketed by its own events)
    per = []
    for i in range(args.steps):
        per.append(event_ms(torch, lambda j: step(i), 1))
    p50, p99 = float(np.percentile(per, 50)), float(np.percentile(per, 99))
    rhs_s = nrhs / (ms_step * 1e-3)
    plain_ms = ms_apply + ms_solve64
    plain_rhs_s = nrhs / (plain_ms * 1e-3)
    log(f"step (apply + mixed solve, one call): {ms_step * 1e3:.0f} us -> {rhs_s:.0f} RHS/s "
        f"(as two calls {ms_step_apart * 1e3:.0f} us); "
        f"plain fp64 {plain_ms * 1e3:.0f} us -> {plain_rhs_s:.0f} RHS/s")

    # --- the drop-in entry points on a host tree -----------------------------
    # dplrgepmv / dplrgepdmm take the reference's host tree: every call
    # flattens it, uploads the level arrays, runs the kernels and copies the
    # result back (no plan survives the call)
    tree0 = D.PlrTree.random(m, args.nlvl, args.rank, 4.0, 1)
    xh = np.random.default_rng(2).standard_normal(m)
    Bh = np.random.default_rng(3).standard_normal((m, nrhs))
    D.gepmv("N", tree0, xh)
    dropin = {}

    def geinv_once():
        Binv, info = D.geinv(tree0, 0.0)
        Binv.free()

    def cpu_sample_once():
        # the CPU baseline's sample through the drop-in: apply, explicit
        # inverse, inverse applied
        D.gepdmm("N", "N", tree0, Bh, nrhs)
        Binv, info = D.geinv(tree0, 0.0)
        D.gepdmm("N", "N", Binv, Bh, nrhs)
        Binv.free()

    for name, fn in (("gepmv_ms", lambda: D.gepmv("N", tree0, xh)),
                     ("gepdmm_ms", lambda: D.gepdmm("N", "N", tree0, Bh, nrhs)),
                     ("geinv_ms", geinv_once),
                     ("cpu_sample_ms", cpu_sample_once)):
        ts = []
        for _ in range(3):
            t1 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t1) * 1e3)
        dropin[name] = float(np.median(ts))
    tree0.free()
    dropin["note"] = (f"host tree in, host result out, median of 3 calls at N={m}: flatten, "
                      f"upload of {abytes / 1e6:.0f} MB, kernels, download; the device-resident "
                      "plan (dplrx_*) is what the other figures time; cpu_sample_ms is the CPU "
                      "baseline's sample (dplrgepdmm, dplrgeinv at eps 0, dplrgepdmm of the "
                      "inverse) through the same entry points")
    log(f"drop-in calls on the host tree: dplrgepmv {dropin['gepmv_ms']:.1f} ms, "
        f"dplrgepdmm ({nrhs} RHS) {dropin['gepdmm_ms']:.1f} ms, dplrgeinv {dropin['geinv_ms']:.1f} ms, "
        f"the CPU baseline's sample {dropin['cpu_sample_ms']:.1f} ms")

    batch = None
    if args.batch_count > 0:
        for P in plans:
            P.close()
        plans = []
        torch.cuda.empty_cache()
        batch = batch_leg(args, torch, D, log)

    precision = precision_leg(args, torch, D, log) if args.precision_legs else None

    grad = None
    if args.grad_m > 0:
        grad = grad_leg(args, torch, D, log)

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
        "ms_per_step_p50": p50,
        "ms_per_step_p99": p99,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f64 apply + {args.lp}/f32-accum refined solve",
        "data": "synthetic",
        "config": {"workload": f"apply + solve, {nrhs} RHS, N={m}, L={args.nlvl}, leaf "
                               f"{m >> args.nlvl}, rank {args.rank}, {args.nmat} matrices rotated",
                   "global_batch": nrhs},
        "dropin": dropin,
        "plain_fp64": {"rhs_s": plain_rhs_s, "ms_per_step": plain_ms,
                       "apply_ms": ms_apply, "solve_ms": ms_solve64},
        "mixed": {"rhs_s": rhs_s, "call": "dplrx_apply_solve_refined_tol",
                  "ms_per_step_two_calls": ms_step_apart,
                  "step_rel_err_vs_fp64": step_err,
                  "solve_ms": ms_mixed, "refine_solves": max(iters),
                  "error_estimate": est, "rel_err_vs_fp64_solve": err, "target_met": target_met,
                  "lowprec": args.lp, "stops_on": "its own error estimate <= 1e-10",
                  "solve_ms_fixed_solves_no_readback": ms_fixed},
        "target_met": target_met,
        "speedup_vs_plain_fp64": plain_ms / ms_step,
        "factor_ms": {"fp64": f64_ms, args.lp: lp_ms, "fp64_again": warm["fp64"],
                      f"{args.lp}_again": warm[args.lp]},
        "incl_factor": {"mixed_rhs_s": nrhs / ((ms_step + lp_ms) * 1e-3),
                        "plain_fp64_rhs_s": nrhs / ((plain_ms + f64_ms) * 1e-3)},
        "apply_fp64_tflops": flops / (ms_apply * 1e-3) / 1e12,
        # the step's FP64 products (one apply, two refinement residuals) on
        # the FP64 MFMA roofline; the matvec above is the HBM-bound kernel
        "roofline_apply": {"bound": "mfma", "achieved": flops / (ms_apply * 1e-3) / 1e12,
                           "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                           "frac": flops / (ms_apply * 1e-3) / 1e12 / FP64_PEAK_TFS,
                           "kernel": "64-RHS FP64 product (k_mm_reduce_fused2 + k_mm_finish<true> "
                                     "+ k_mm_expand3)",
                           "flops_per_launch": flops, "us_per_launch": ms_apply * 1e3},
        # one low-precision solve of the refinement (the bulk of the step)
        "roofline_mixed": {"bound": "hbm", "achieved": lp_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": lp_gbs / HBM_PEAK_GBS,
                           "traffic": measured_traffic(f"mixed_solve_{args.lp}_N65536_L10_r32_64rhs.md",
                                                       "traffic_bytes_per_solve")
                           if m == 65536 and nrhs == 64 else None,
                           "kernel": f"one {args.lp} solve ("
                                     + ("k_lp_colmax + k_lp_tail + k_lp_up_t0 + 7 x k_up_step + "
                                        "k_lp_upx)" if args.lp == "f16x2" else
                                        "k_lp_colmax + k_lp_tail + 7 x (k_lp_level + "
                                        "k_lp_finish1))"),
                           "bytes_per_launch": lp_bytes, "us_per_launch": ms_one * 1e3},
        "roofline": {"bound": "hbm", "achieved": mv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": mv_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "single-RHS FP64 matvec (k_mv_reduce3 + k_mv_expand4)",
                     "bytes_per_launch": mv_bytes, "us_per_launch": ms_mv * 1e3},
        "cpu_baseline": cpu,
        "batch": batch,
        "precision": precision,
        "gradient": grad,
        "wall_s_timed": wall,
    }
    if not target_met:
        out["value"] = None
    print(json.dumps(out), flush=True)
    for p in plans:
        p.close()
    if not target_met:
        sys.exit(1)


if __name__ == "__main__":
    main()
