"""Benchmark of the single-GPU AMG-FCG path on MI355X.

Workload (one step): the 27-point Laplacian on a 128^3 grid (2,097,152 rows,
55,742,968 nonzeros) with 32 right-hand sides drawn N(0,1) (numpy PCG64,
seed 3), solved to relative residual 1e-8 by FCG preconditioned with one
BCMG V-cycle (3 pairwise matching sweeps per level, 1 pre / 1 post l1-Jacobi
sweep, 20 coarsest sweeps), FP32 hierarchy and V-cycle under the FP64 FCG,
all 32 columns in one batched solve.  A step is AMG setup + batched solve
(time-to-solution); inputs are resident in HBM when the timed region starts.

Reported beside it: the plain FP64 path solving the 32 right-hand sides one at
a time (one setup), the FP64 SpMV of the fine matrix against the HBM roofline
(HIP-event timed on the library stream), a STREAM-copy ceiling, and the CPU
oracle's FP64 setup + solve on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "AMG-FCG time-to-1e-8 (s) + SpMV HBM GB/s, 27p 128^3; mixed/batched vs FP64"
HBM_PEAK_GBS = 8000.0


def spmv_fp64_bytes(n, nnz):
    # 12 nnz (val + col) + 4 (n + 1) (row pointers) + 8 n (x) + 8 n (y)
    return 12 * nnz + 4 * (n + 1) + 16 * n


def smoother_bytes(n, nnz, k, fp32):
    # one l1-Jacobi sweep y = x + (b - A x) / D on k columns: the matrix
    # (val + col + row pointers) once, x, b and y once per column, D once
    w = 4 if fp32 else 8
    return (w + 4) * nnz + 4 * (n + 1) + 3 * w * n * k + w * n


def csr_bytes(n, nnz):
    return 12 * nnz + 4 * (n + 1)


def kernel_rooflines(bx, ctx, A, n, nnz, k, reps):
    """Achieved HBM bandwidth of the smoother (FP64 one column; FP32 k columns,
    the batched V-cycle's fine sweep) and of the Galerkin SpGEMMs of the first
    fine matching sweep (A P, then P^T (A P)), HIP-event timed.  SpGEMM bytes are
    the compulsory ones (both operands and the product once each)."""
    import numpy as np

    out = {}
    for name, prec, kk in (("smoother_fp64_k1", 0, 1), (f"smoother_fp32_k{k}", 1, k)):
        ms = A.time_smoother(precision=prec, k=kk, reps=reps)
        by = smoother_bytes(n, nnz, kk, prec == 1)
        out[name] = {"kernel": "l1-Jacobi sweep (JACOBI epilogue of the autotuned SpMV/SpMM)", "bound": "hbm",
                     "avg_launch_ms": ms, "algorithmic_bytes": by, "achieved": by / (ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    M = A.matching(np.ones(n))
    agg, pval, nc = bx.make_prolongator(ctx, M, np.ones(n))
    P = bx.CSR(ctx, n, nc, np.arange(n + 1, dtype=np.int32), agg, pval)
    R = P.transpose()
    AP = A.spgemm(P)
    for name, L, Rm, nl, nzl, nr, nzr in (("spgemm_AP", A, P, n, nnz, n, n), ("spgemm_RAP", R, AP, nc, n, n, AP.nnz)):
        ms, nzc = L.time_spgemm(Rm, reps=max(1, reps // 4))
        by = csr_bytes(nl, nzl) + csr_bytes(nr, nzr) + csr_bytes(nl, nzc)
        out[name] = {"kernel": "deterministic SpGEMM (count, bin, sort-and-sum rows, compact), whole product",
                     "bound": "hbm", "avg_product_ms": ms, "algorithmic_bytes": by, "nnz_c": nzc,
                     "achieved": by / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    for h in (AP, R, P):
        h.close()
    return out


def pmc_traffic(name="spmv_fp64_lap27_128.txt"):
    """HBM bytes per launch of the fine FP64 SpMV from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def oracle_params(rtol):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from bootcmatchgx_amd import abi

    p = abi.reference_defaults()
    p.solver = abi.SOLVER["FCG"]
    p.preconditioner = abi.PREC["BCMG"]
    p.rtol = rtol
    p.prerelax_sweeps = 1
    p.postrelax_sweeps = 1
    p.dispnorm = 0
    return p


def cpu_baseline(size, nrhs, rtol, samples=5):
    """The oracle (test infrastructure) timed on the host: FP64 setup, then FCG solves
    of the first `samples` columns; value = setup + nrhs x the median solve."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O

    threads = O.lib().or_num_threads()
    A = O.laplacian3d(27, size, size, size)
    n = A.dims()[0]
    Bh = np.random.default_rng(3).normal(size=(n, samples))
    p = oracle_params(rtol)
    t0 = time.perf_counter()
    P = O.Prec(A, p)
    setup = time.perf_counter() - t0
    solves, its = [], []
    for j in range(samples):
        t1 = time.perf_counter()
        _, info = O.fcg(A, P, p, np.ascontiguousarray(Bh[:, j]))
        solves.append(time.perf_counter() - t1)
        its.append(info["niter"])
    med = float(np.median(solves))
    return {
        "value": setup + nrhs * med,
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": f"27p {size}^3 FP64 BCMG setup ({setup:.2f} s) + FCG solves of columns 0..{samples - 1} "
                  f"({', '.join(f'{t:.2f}' for t in solves)} s; {its} iterations); "
                  f"value = setup + {nrhs} x the median solve ({med:.2f} s)",
        "setup_s": setup,
        "solve_s": solves,
    }


def timed_solve(bx, ctx, A, B, X, Z, k, p, prec_params=None):
    """setup + solve of k device-resident columns from a zero guess; seconds and SolverOut."""
    X.copy_from(Z)
    ctx.synchronize()
    t0 = time.perf_counter()
    P = bx.prec_setup(ctx, A, prec_params or p)
    t1 = time.perf_counter()
    out = bx.solve_device(ctx, A, B.ptr, X.ptr, k, p, P)
    t2 = time.perf_counter()
    P.finalize()
    return {"time_s": t2 - t0, "setup_s": t1 - t0, "solve_s": t2 - t1, "niter": out["niter"],
            "rel_res": out["rel_res"]}


def config_measurements(bx, ctx, A128, n128, steps):
    """One line per named configuration besides the headline (each solve from a zero
    guess to relative residual 1e-8, inputs resident in HBM, one untimed warm-up run
    that also tunes the kernels, then the mean of `steps` timed runs)."""
    import numpy as np

    def mean_runs(run):
        run()
        rs = [run() for _ in range(max(1, steps))]
        out = dict(rs[-1])
        for key in ("time_s", "setup_s", "solve_s"):
            out[key] = float(np.mean([r[key] for r in rs]))
        return out

    res = {}
    common = dict(solver="FCG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, dispnorm=0)

    # C0: 7-point 50^3, b = A x*, x* ~ U[-1,1] (seed 0), FP64 BCMG; the CPU oracle then the GPU
    n, row, col, val = bx.laplacian3d(7, 50, 50, 50)
    A = bx.CSR(ctx, n, n, row, col, val)
    xs = np.random.default_rng(0).uniform(-1.0, 1.0, size=n)
    XS, B, X, Z = (bx.DeviceArray(ctx, (n,)) for _ in range(4))
    XS.upload(xs)
    Z.upload(np.zeros(n))
    A.spmm_device(XS.ptr, B.ptr, 1, n)
    p = bx.Params(preconditioner="BCMG", precision="FP64", **common)
    gpu = mean_runs(lambda: timed_solve(bx, ctx, A, B, X, Z, 1, p))
    gpu["err_vs_xstar"] = float(np.linalg.norm(X.download() - xs) / np.linalg.norm(xs))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O

    oa = O.laplacian3d(7, 50, 50, 50)
    b = B.download()
    t0 = time.perf_counter()
    op = O.Prec(oa, oracle_params(1e-8))
    t1 = time.perf_counter()
    xo, oi = O.fcg(oa, op, oracle_params(1e-8), b)
    t2 = time.perf_counter()
    res["c0_lap7_50"] = {"workload": "7-point Laplacian 50^3, b = A x*, x* U[-1,1] seed 0, FCG + FP64 BCMG V(1,1)",
                         "gpu": gpu, "cpu_oracle": {"time_s": t2 - t0, "setup_s": t1 - t0, "solve_s": t2 - t1,
                                                    "niter": oi["niter"], "cores": O.lib().or_num_threads(),
                                                    "err_vs_xstar": float(np.linalg.norm(xo - xs) / np.linalg.norm(xs))}}
    for h in (XS, B, X, Z, A):
        h.close()

    # C1: 27-point 128^3, b = A x*, x* ~ U[-1,1] (seed 1), FP64: BCMG and AFSAI
    n = n128
    xs = np.random.default_rng(1).uniform(-1.0, 1.0, size=n)
    XS, B, X, Z = (bx.DeviceArray(ctx, (n,)) for _ in range(4))
    XS.upload(xs)
    Z.upload(np.zeros(n))
    A128.spmm_device(XS.ptr, B.ptr, 1, n)
    c1 = {"workload": "27-point Laplacian 128^3, b = A x*, x* U[-1,1] seed 1, FCG, all FP64"}
    for prec in ("BCMG", "AFSAI"):
        p = bx.Params(preconditioner=prec, precision="FP64", **common)
        r = mean_runs(lambda: timed_solve(bx, ctx, A128, B, X, Z, 1, p))
        r["err_vs_xstar"] = float(np.linalg.norm(X.download() - xs) / np.linalg.norm(xs))
        c1[prec.lower()] = r
    res["c1_lap27_128"] = c1
    for h in (XS, B, X, Z):
        h.close()

    # C2: 7-point diffusion 160^3, conductivities exp(2 g) (seed 2), b ~ U[-1,1] (seed 2):
    # FP32 hierarchy + V-cycle vs FP64, both under the FP64 FCG
    n, row, col, val = bx.diffusion7p(160, 160, 160, 0, 2.0, 0.0, 2)
    A = bx.CSR(ctx, n, n, row, col, val)
    del row, col, val
    B, X, Z = (bx.DeviceArray(ctx, (n,)) for _ in range(3))
    B.upload(np.random.default_rng(2).uniform(-1.0, 1.0, size=n))
    Z.upload(np.zeros(n))
    c2 = {"workload": "7-point diffusion 160^3, face conductivity exp(2 g) seed 2, b U[-1,1] seed 2, FCG + BCMG V(1,1)"}
    sol = {}
    for precision in ("FP32", "FP64"):
        p = bx.Params(preconditioner="BCMG", precision=precision, **common)
        c2[precision.lower()] = mean_runs(lambda: timed_solve(bx, ctx, A, B, X, Z, 1, p))
        sol[precision] = X.download()
    c2["fp32_vs_fp64_solution_rel_diff"] = float(np.linalg.norm(sol["FP32"] - sol["FP64"]) / np.linalg.norm(sol["FP64"]))
    res["c2_diff7_160"] = c2
    for h in (B, X, Z, A):
        h.close()

    # C3: the 27-point 128^3 matrix, k = 1, 8, 32, 64 columns N(0,1) (seed 3), FP32 hierarchy, batched
    c3 = {"workload": "27-point Laplacian 128^3, k columns N(0,1) seed 3, FCG + FP32 BCMG V(1,1), one batched solve"}
    for k in (1, 8, 32, 64):
        n = n128
        B, X, Z = (bx.DeviceArray(ctx, (n, k)) for _ in range(3))
        B.upload(np.random.default_rng(3).normal(size=(n, k)).ravel(order="F"))
        Z.upload(np.zeros(n * k))
        p = bx.Params(preconditioner="BCMG", precision="FP32", **common)
        r = mean_runs(lambda: timed_solve(bx, ctx, A128, B, X, Z, k, p))
        r["s_per_column"] = r["time_s"] / k
        c3[f"k{k}"] = r
        for h in (B, X, Z):
            h.close()
    res["c3_lap27_128_batched"] = c3
    return res


def adjoint_timing(ctx, N=128):
    """Forward solve + adjoint solve through the torch.autograd shim on the 7-point
    diffusion operator with per-face coefficients U[0.5, 2] (seed 4), loss
    0.5 ||x - x_target||^2 (x_target ~ N(0,1) seed 5, b ~ N(0,1) seed 6), FP64
    BCMG hierarchy built once and reused by the adjoint solve.  CPU tensors: the
    times include the host<->device copies of x, b and the gradients."""
    import numpy as np

    import bootcmatchgx_amd as bx
    from bootcmatchgx_amd import autograd as AG

    torch = AG.torch
    n, row, col, val = bx.diffusion7p(N, N, N, 1, 0.5, 2.0, 4)
    xt = torch.from_numpy(np.random.default_rng(5).normal(size=n))
    b = np.random.default_rng(6).normal(size=n)
    p = bx.Params(solver="FCG", preconditioner="BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, dispnorm=0)
    t0 = time.perf_counter()
    system = AG.SPDSystem(n, row, col, val, p, ctx=ctx)
    t1 = time.perf_counter()
    v = torch.tensor(val, requires_grad=True)
    bt = torch.tensor(b, requires_grad=True)
    x = AG.sparse_solve(v, bt, system)
    it_f = system.last_info["niter"]
    L = 0.5 * ((x - xt) ** 2).sum()
    t2 = time.perf_counter()
    L.backward()
    t3 = time.perf_counter()
    it_a = system.last_info["niter"]
    system.close()
    return {"workload": f"7-point diffusion U[0.5,2] {N}^3, FP64 BCMG, rtol 1e-8", "n": n, "nnz": len(col),
            "setup_s": t1 - t0, "forward_s": t2 - t1, "adjoint_s": t3 - t2, "niter_forward": it_f,
            "niter_adjoint": it_a}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--nrhs", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 one-at-a-time comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the autograd forward + adjoint timing")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations' measurements")
    ap.add_argument("--spmv-only", type=int, default=0, metavar="REPS",
                    help="only time REPS launches of the FP64 SpMV of the fine matrix (profiling)")
    ap.add_argument("--kernels-only", type=int, default=0, metavar="REPS",
                    help="only time the smoother and the Galerkin SpGEMMs of the fine matrix (profiling)")
    ap.add_argument("--solve-only", action="store_true",
                    help="only one setup and one batched solve of the headline workload (profiling)")
    ap.add_argument("--afsai-only", type=int, default=0, metavar="REPS",
                    help="only REPS AFSAI setups of the 27-point matrix at --size (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU row-block path is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import bootcmatchgx_amd as bx

    N, k, rtol = args.size, args.nrhs, 1e-8
    ctx = bx.Context(0)
    n, row, col, val = bx.laplacian3d(27, N, N, N)
    nnz = len(col)
    A = bx.CSR(ctx, n, n, row, col, val)
    del row, col, val
    # the pattern's row-tile index, built once per matrix and kept with it (the first setup
    # would build it otherwise): its cost is reported apart
    tile_s = A.analyse()
    if args.spmv_only:
        ms = A.time_spmv(precision=0, k=1, reps=args.spmv_only)
        gbs = spmv_fp64_bytes(n, nnz) / (ms * 1e-3) / 1e9
        copy_gbs = ctx.stream_copy_gbs(1 << 30, 20)  # k_copy16 over 1 GiB: counter calibration
        print(json.dumps({"kernel": "spmv fp64", "n": n, "nnz": nnz, "avg_launch_ms": ms, "gbs": gbs,
                          "algorithmic_bytes": spmv_fp64_bytes(n, nnz), "copy_bytes_per_launch": 2 * (1 << 30),
                          "copy_gbs": copy_gbs}))
        return
    if args.kernels_only:
        res = kernel_rooflines(bx, ctx, A, n, nnz, k, args.kernels_only)
        res["copy_gbs"] = ctx.stream_copy_gbs(1 << 30, 20)  # k_copy16 over 1 GiB: counter calibration
        print(json.dumps(res))
        return
    if args.afsai_only:
        p = bx.Params(solver="FCG", preconditioner="AFSAI", rtol=rtol, dispnorm=0)
        times = []
        for _ in range(args.afsai_only):
            ctx.synchronize()
            t0 = time.perf_counter()
            P = bx.prec_setup(ctx, A, p)
            times.append(time.perf_counter() - t0)
            P.finalize()
        print(json.dumps({"afsai_setup_s": times, "n": n, "nnz": nnz}))
        return
    if args.solve_only:
        B = bx.DeviceArray(ctx, (n, k))
        B.upload(np.random.default_rng(3).normal(size=(n, k)).ravel(order="F"))
        X = bx.DeviceArray(ctx, (n, k))
        X.upload(np.zeros(n * k))
        p = bx.Params(solver="FCG", preconditioner="BCMG", rtol=rtol, prerelax_sweeps=1, postrelax_sweeps=1,
                      dispnorm=0, precision="FP32")
        ctx.synchronize()
        t0 = time.perf_counter()
        P = bx.prec_setup(ctx, A, p)
        t1 = time.perf_counter()
        out = bx.solve_device(ctx, A, B.ptr, X.ptr, k, p, P)
        t2 = time.perf_counter()
        P.finalize()
        print(json.dumps({"setup_s": t1 - t0, "solve_s": t2 - t1, "niter": out["niter"], "rel_res": out["rel_res"],
                          "note": "first setup and solve of the process: includes the one-time autotuning"}))
        return
    Bh = np.asfortranarray(np.random.default_rng(3).normal(size=(n, k)))
    B = bx.DeviceArray(ctx, (n, k))
    B.upload(Bh.ravel(order="F"))
    X = bx.DeviceArray(ctx, (n, k))
    Z = bx.DeviceArray(ctx, (n, k))
    Z.upload(np.zeros(n * k))

    common = dict(solver="FCG", preconditioner="BCMG", rtol=rtol, prerelax_sweeps=1, postrelax_sweeps=1, dispnorm=0)
    p32 = bx.Params(precision="FP32", **common)
    p64 = bx.Params(precision="FP64", **common)

    def step():
        X.copy_from(Z)
        ctx.synchronize()  # the initial guess is reset before the clock starts
        t0 = time.perf_counter()
        P = bx.prec_setup(ctx, A, p32)
        t1 = time.perf_counter()
        out = bx.solve_device(ctx, A, B.ptr, X.ptr, k, p32, P)
        t2 = time.perf_counter()
        P.finalize()
        return t2 - t0, t1 - t0, t2 - t1, out

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    times, setups, solves, outs = [], [], [], []
    for _ in range(args.steps):
        t, ts, tv, out = step()
        times.append(t)
        setups.append(ts)
        solves.append(tv)
        outs.append(out)
    ctx.synchronize()
    X32 = X.download().reshape((n, k), order="F")
    value = float(np.mean(times))
    last = outs[-1]

    # plain FP64 path, one right-hand side at a time (one setup)
    fp64 = None
    if not args.no_fp64:
        X.copy_from(Z)
        ctx.synchronize()  # the initial guess is reset before the clock starts
        t0 = time.perf_counter()
        P = bx.prec_setup(ctx, A, p64)
        t1 = time.perf_counter()
        it64 = []
        for j in range(k):
            o = bx.solve_device(ctx, A, B.at(j * n), X.at(j * n), 1, p64, P)
            it64.append(o["niter"])
        t2 = time.perf_counter()
        P.finalize()
        X64 = X.download().reshape((n, k), order="F")
        fp64 = {"time_s": t2 - t0, "setup_s": t1 - t0, "solve_s": t2 - t1, "niter_max": max(it64),
                "err_mixed_vs_fp64": float(np.linalg.norm(X32 - X64) / np.linalg.norm(X64))}
        # the same FP64 hierarchy with the k columns batched: batching without mixed
        # precision (the second of two runs: the first tunes the batched kernels)
        for _ in range(2):
            X.copy_from(Z)
            ctx.synchronize()
            t0 = time.perf_counter()
            P = bx.prec_setup(ctx, A, p64)
            t1 = time.perf_counter()
            o = bx.solve_device(ctx, A, B.ptr, X.ptr, k, p64, P)
            t2 = time.perf_counter()
            P.finalize()
        fp64["batched"] = {"time_s": t2 - t0, "setup_s": t1 - t0, "solve_s": t2 - t1, "niter": o["niter"]}

    # SpMV FP64 of the fine matrix against the HBM roofline (HIP events on the library stream)
    spmv_ms = A.time_spmv(precision=0, k=1, reps=50)
    spmv_bytes = spmv_fp64_bytes(n, nnz)
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    spmm32_ms = A.time_spmv(precision=1, k=k, reps=10)
    copy_gbs = ctx.stream_copy_gbs(1 << 30, 20)
    kernels = kernel_rooflines(bx, ctx, A, n, nnz, k, 20)

    adjoint = None if args.no_adjoint else adjoint_timing(ctx)
    configs = None if args.no_configs else config_measurements(bx, ctx, A, n, 1)

    cpu = None if args.no_cpu else cpu_baseline(N, k, rtol)

    res = {
        "metric": METRIC,
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": value * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 hierarchy / f64 krylov",
        "data": "synthetic",
        "config": {
            "workload": f"27-point Laplacian {N}^3, {k} RHS N(0,1) seed 3, FCG+BCMG rtol 1e-8, "
                        "V-cycle 1/1 l1-Jacobi, FP32 hierarchy, batched",
            "n": n, "nnz": nnz, "nrhs": k,
        },
        "setup_s": float(np.mean(setups)),
        "solve_s": float(np.mean(solves)),
        "niter": last["niter"],
        "final_rel_res": last["rel_res"],
        "fp64_one_at_a_time": fp64,
        "speedup_vs_fp64_one_at_a_time": (fp64["time_s"] / value) if fp64 else None,
        "spmv_fp64_gbs": spmv_gbs,
        "spmm_fp32_k32_ms": spmm32_ms,
        "stream_copy_gbs": copy_gbs,
        "autograd_forward_adjoint": adjoint,
        "configs": configs,
        "roofline": {
            "kernel": "autotuned SpMV of the fine 27-point matrix, FP64 (the row-tile kernel k_spmv_tiled when the "
                      "tuner picks it: x of each 64-row tile's distinct columns staged in LDS, 2-byte column positions)",
            "tile_index_s": tile_s,
            "note": "achieved = CSR algorithmic bytes / time; the row-tile kernel moves fewer bytes than CSR (traffic)",
            "bound": "hbm",
            "achieved": spmv_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": spmv_gbs / HBM_PEAK_GBS,
            "traffic": pmc_traffic() if N == 128 else None,
            "traffic_source": "profiles/spmv_fp64_lap27_128.txt (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
            "algorithmic_bytes": spmv_bytes,
            "avg_launch_ms": spmv_ms,
        },
        "kernel_rooflines": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))
    A.close()
    ctx.close()


if __name__ == "__main__":
    main()
