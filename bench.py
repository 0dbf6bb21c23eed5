#!/usr/bin/env python3
"""Benchmark of the hot path on one MI355X.

Workload (3-D linear elasticity, the configuration the headline metric is
quoted on): unit cube, 64^3 trilinear hexes (274,625 nodes, 823,875 dofs),
2x2x2 Gauss, E = 1, nu = 0.3, body force (0,0,-1), clamped x = 0 face,
3x3-block BCSR.  One step = assembly of the tangent and load into the fixed
pattern + Dirichlet + Jacobi-PCG to ||b - A x|| / ||b|| <= 1e-8 with FP32
inner iterations and FP64 refinement.  Mesh, DoF numbering and the pattern are
built once (device resident) before timing.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-serve] [--no-traffic]

Prints one JSON line.  `value` is solves/s; the same line carries p50/p99
step latency, the plain FP64 GPU paths (the correctness baseline), the
roofline of the CG SpMV (its HBM traffic measured by two rocprofv3 counter
runs of this script's traffic probe, started first as child processes), a
bounded served stream of the mixed Poisson / elasticity requests (solves/s,
p50/p99 at three arrival rates, against the FP64 path one request at a
time), and the reference CPU path timed on this host with the GPU
solutions' error against its solve.  --workload selects the other modes.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
TOL = 1e-8


def bcsr_bytes(nb, nbr, n, s, ib=4):
    """BCSR3 SpMV bytes: blocks of 9 values and one column index of ib bytes
    (2 with 16-bit column offsets), the row structure, x read and y written."""
    return nb * (9 * s + ib) + (nbr + 1) * 4 + 2 * n * s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """CPU model (from /proc/cpuinfo) and the threads this process may use:
    its affinity set, capped by OMP_NUM_THREADS when that is set."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return model or platform.processor() or platform.machine(), threads


def bcsr_to_csr(ptr, col, val):
    """The scalar CSR of a BCSR3 matrix in the reference layout (block values
    row-major): row 3I+i holds, block by block, columns 3J..3J+2."""
    nbr = len(ptr) - 1
    blen = np.diff(ptr)
    brk = np.repeat(np.arange(nbr, dtype=np.int64), blen)
    koff = np.arange(len(col), dtype=np.int64) - ptr[brk]
    base = 9 * ptr[brk] + 3 * koff
    i = np.arange(3, dtype=np.int64)
    pos = (base[:, None, None] + (3 * i[None, :] * blen[brk][:, None])[:, :, None]
           + i[None, None, :]).ravel()
    ccol = np.empty(9 * len(col), np.int64)
    cval = np.empty(9 * len(col))
    ccol[pos] = np.broadcast_to((3 * col)[:, None, None] + i[None, None, :], (len(col), 3, 3)).ravel()
    cval[pos] = val
    rptr = np.zeros(3 * nbr + 1, np.int64)
    rptr[1:] = np.cumsum(np.repeat(3 * blen, 3))
    return rptr, ccol, cval


def cpu_baseline(w, xs, fp64_cg_iters):
    """The reference's own CPU path on this host, beside the GPU run.

    Solve: the reference's cg_solve (cg_solve.hpp:33-81) with its serial
    bcsr_spmv (bcsr_spmv.hpp:23-62), compiled from the reference sources with
    its Release flags, run once in full on the assembled system to
    ||r||^2/||b||^2 < 1e-16, i.e. ||r||/||b|| < 1e-8 (the reference's BCSR
    path has no threads).  Its solution is the one the GPU solutions in `xs`
    are compared with.  Beside it, the reference's threaded path: cg_solve's
    loop body with csr_spmv (OpenMP schedule(static,128), csr_spmv.hpp:36-38)
    on the same matrix in CSR, timed for a bounded number of iterations.
    Assembly: CGAssembly's OpenMP element loop (the oracle's restatement) and
    the reference's sparse_bcsr triplet constructor on a 64 x 64 x 4 slab
    (1/16 of the elements), scaled by 16, on all threads and on one."""
    from oracle.oracle import Reference, Oracle, physics
    from yafel_amd import workloads
    if not Reference.available(fast=True):
        return None
    model, nthreads = host_cpu()
    ref = Reference(fast=True)
    ptr, col, val = w.A.arrays()
    b = w.b.download()
    M = ref.matrix_from_arrays(3, w.n, ptr, col, val)
    t = time.perf_counter()
    x_ref = M.cg(b, tol=TOL * TOL)
    solve_s = time.perf_counter() - t
    relres = float(np.linalg.norm(b - M.spmv(x_ref)) / np.linalg.norm(b))
    nref = float(np.linalg.norm(x_ref))
    errs = {k: float(np.linalg.norm(x - x_ref) / nref) for k, x in xs.items()}
    del M
    log(f"cpu: reference cg_solve (BCSR, serial) {solve_s:.2f} s, rel residual {relres:.2e}; errors {errs}")
    # the threaded solve path: csr_spmv with OpenMP
    rptr, ccol, cval = bcsr_to_csr(ptr, col, val)
    del val
    Mc = ref.matrix_from_arrays(1, w.n, rptr, ccol, cval)
    del rptr, ccol, cval
    t = time.perf_counter()
    Mc.cg_iterations(b, 2)
    t2 = (time.perf_counter() - t) / 2
    k = int(max(3, min(100, 6.0 / max(t2, 1e-6))))
    t = time.perf_counter()
    Mc.cg_iterations(b, k)
    csr_iter = (time.perf_counter() - t) / k
    del Mc
    try:
        import ctypes
        omp_threads = ctypes.CDLL("libgomp.so.1").omp_get_max_threads()
    except OSError:
        omp_threads = None
    # assembly sample
    orc = Oracle(fast=True)
    nodes, cells = orc.structured_mesh(3, 64, 64, 4, 0)
    el, dn = orc.dofmanager(3, 1, nodes, cells)
    ph = physics(1, 3, lam=workloads.LAMBDA, mu=workloads.MU, body=(0, 0, -1))
    asm = {}
    for nt in sorted({1, nthreads}):
        t = time.perf_counter()
        rows, cols, vals, _ = orc.assemble_threads(nt, 3, 1, 3, dn, el, ph)
        t_el = time.perf_counter() - t
        t = time.perf_counter()
        B = ref.matrix(rows, cols, vals, block=3)
        t_tr = time.perf_counter() - t
        del B, rows, cols, vals
        asm[nt] = {"element_loop_s": 16 * t_el, "triplets_to_bcsr_s": 16 * t_tr, "total_s": 16 * (t_el + t_tr)}
    csr_solve_s = csr_iter * fp64_cg_iters
    best_solve = min(solve_s, csr_solve_s)
    return {
        "value": 1.0 / (best_solve + asm[nthreads]["total_s"]), "unit": "solves/s", "cores": nthreads,
        "kind": "reference",
        "sample": (f"reference cg_solve + serial bcsr_spmv run once in full on the assembled 64^3 system "
                   f"({solve_s:.1f} s, rel residual {relres:.1e}); reference cg_solve loop body with OpenMP "
                   f"csr_spmv on the same matrix in CSR, {k} iterations ({csr_iter * 1e3:.1f} ms/iter) scaled by "
                   f"the {fp64_cg_iters} FP64 CG iterations; CGAssembly element loop (OpenMP, oracle restatement) "
                   f"+ reference sparse_bcsr(triplets) on a 64x64x4 slab scaled x16; value = best solve + "
                   f"assembly on {nthreads} threads; -O3 -ffast-math"),
        "cpu": model, "threads": nthreads, "omp_threads": omp_threads,
        "solve_bcsr_serial_s": solve_s, "solve_bcsr_serial_rel_residual": relres,
        "solve_csr_openmp_ms_per_iter": csr_iter * 1e3, "solve_csr_openmp_s": csr_solve_s,
        "assembly": {str(k_): v for k_, v in asm.items()},
        "value_1_thread": 1.0 / (solve_s + asm[1]["total_s"]),
        "gpu_solution_rel_error_vs_reference": errs,
    }


KERNEL_A = "k_spmv_fused<3, 4, float, OpCgQ<float>>"


def _counter_rows(d):
    """(kernel name, per-dispatch counter value) rows of the counter CSVs under d."""
    import csv
    import glob
    per = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                key = (f, row["Dispatch_Id"])
                name, v = row["Kernel_Name"], float(row["Counter_Value"])
                per[key] = (name, per.get(key, (name, 0.0))[1] + v)
    return list(per.values())


def measured_traffic(n):
    """HBM bytes per launch of the CG SpMV kernel, measured now: two child runs
    of this script's traffic probe under rocprofv3 --pmc, FETCH_SIZE then
    WRITE_SIZE, each scaled by what the same counter reads for the copy kernel
    moving a known 1 GiB (read and written).  Only launches that ran the
    product count (at least half the largest launch's fetch).  Runs before
    this process touches the GPU; returns (bytes or None, detail)."""
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if not exe:
        return None, {"error": "rocprofv3 not found"}
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="yafel_pmc_", dir="/tmp")
        cmd = [exe, "--pmc", counter, "--kernel-include-regex", "k_spmv_fused|k_copy", "--output-format", "csv",
               "-d", d, "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--workload",
               "traffic-probe", "--n", str(n)]
        t = time.perf_counter()
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=240)
            out, rc = r.stdout, r.returncode
        except subprocess.TimeoutExpired as e:
            out, rc = (e.stdout or b"").decode(errors="replace") if isinstance(e.stdout, bytes) else (e.stdout or ""), -9
        rows = _counter_rows(d)
        shutil.rmtree(d, ignore_errors=True)
        log(f"traffic {counter}: rc {rc}, {len(rows)} dispatches, {time.perf_counter() - t:.1f} s")
        # the profiler may end with a fault of its own after the program's last
        # line; the probe's completion line and whole results are what count
        if "traffic-probe done" not in out or not rows:
            return None, {"error": f"{counter} run rc {rc}: {out[-400:]}"}
        cg = [v for name, v in rows if "k_spmv_fused" in name and "OpCgQ<float>" in name]
        cp = [v for name, v in rows if "k_copy" in name]
        if not cg or not cp:
            return None, {"error": f"{counter}: kernels not found in the counter output"}
        full = [v for v in cg if v >= 0.5 * max(cg)]
        res[counter] = {"kib_per_launch": float(np.mean(full)), "launches": len(full), "skipped": len(cg) - len(full),
                        "copy_kib": float(np.median(cp)), "factor": (1 << 30) / (float(np.median(cp)) * 1024)}
    fetch = res["FETCH_SIZE"]["kib_per_launch"] * 1024 * res["FETCH_SIZE"]["factor"]
    write = res["WRITE_SIZE"]["kib_per_launch"] * 1024 * res["WRITE_SIZE"]["factor"]
    return fetch + write, {"read_bytes": fetch, "write_bytes": write, **{k.lower(): v for k, v in res.items()}}


def traffic_probe(args):
    """The command the counter runs profile: one mixed c2 step (the CG SpMV
    launches) and three launches of the 1 GiB copy kernel (the calibration)."""
    from yafel_amd import hip, workloads
    ctx = hip.Context(0)
    w = workloads.c2(ctx, n=args.n)
    info = w.step(method=hip.PCG_JACOBI, precision=hip.MIXED, stop=hip.STOP_RELRES, tol=TOL)
    ctx.copy_bandwidth(1 << 30, 2)
    ctx.sync()
    print(f"traffic-probe done: {info.iterations} iterations", flush=True)


def req_class(m):
    return "3d_q1_elasticity" if m["dim"] == 3 else f"2d_q{m['p']}_poisson"


def run_stream(hip, ctx, mats, arrivals, budget, slots, setup_threads=0, graphs=False, cg_passes=0,
               two_pass_limit=0):
    """Feed requests at their arrival times into a server; returns per-request
    latency (arrival -> found finished), makespan, the results, the server
    statistics and each request's (admission wait, setup, wait to join,
    solve) times in ms.  graphs: replay chunks from HIP graphs;
    cg_passes, two_pass_limit: kernels per batched CG iteration (0, 0: the
    library's default)."""
    srv = hip.Server(ctx, budget, slots, setup_threads)
    if graphs:
        srv.set_graphs(True)
    if cg_passes or two_pass_limit:
        srv.set_cg_passes(cg_passes, two_pass_limit)
    n = len(mats)
    t0 = time.perf_counter()
    i, done_t, rid_of, times = 0, {}, {}, {}
    while len(done_t) < n:
        now = time.perf_counter() - t0
        while i < n and arrivals[i] <= now:
            m = mats[i]
            rid = srv.submit(m["dim"], m["p"], m["dpn"], m["nodes"], m["cells"], m["ph"], load=m["load"],
                             bc=m["bc"], tol=TOL)
            rid_of[rid] = i
            i += 1
        fin = srv.progress()
        tnow = time.perf_counter() - t0
        for rid in fin:
            done_t[rid_of[rid]] = tnow
            times[rid_of[rid]] = srv.times(rid)
        if not fin and srv.stats()["active"] == 0 and i < n and arrivals[i] > tnow:
            time.sleep(min(arrivals[i] - tnow, 0.002))
    makespan = max(done_t.values())
    results = {}
    for rid, j in rid_of.items():
        results[j] = srv.result(rid)
    st = srv.stats()
    srv.close()
    lat = np.array([done_t[j] - arrivals[j] for j in range(n)])
    return lat, makespan, results, st, times


def latency_breakdown(mats, lat, times):
    """Median (and p99 of the total) per request class: admission wait, setup,
    wait to join the batch, solve, and arrival -> result."""
    out = {}
    for cls in sorted({req_class(m) for m in mats}):
        idx = [j for j, m in enumerate(mats) if req_class(m) == cls]
        t = np.array([times[j] for j in idx])
        out[cls] = {"requests": len(idx), "wait_ms_p50": float(np.median(t[:, 0])),
                    "setup_ms_p50": float(np.median(t[:, 1])), "join_ms_p50": float(np.median(t[:, 2])),
                    "solve_ms_p50": float(np.median(t[:, 3])), "total_ms_p50": float(np.median(lat[idx]) * 1e3),
                    "total_ms_p99": float(np.percentile(lat[idx], 99) * 1e3)}
    return out


def fp64_one_at_a_time(hip, ctx, mats):
    """The plain FP64 GPU path (FP64 assembly, Jacobi-PCG in FP64 to the same
    relative residual), one request at a time; returns solves/s."""
    t = time.perf_counter()
    for m in mats:
        mesh = hip.Mesh(ctx, m["dim"], m["p"], m["dpn"], m["nodes"], m["cells"])
        A = mesh.pattern(hip.BCSR3 if m["dpn"] == 3 else hip.CSR, hip.FP64)
        mesh.assemble(m["ph"], A, None, hip.TANGENT)
        b = ctx.upload(m["load"])
        mask = ctx.upload(np.zeros(mesh.ndofs, np.uint8))
        if m["bc"] == hip.BC_BOX:
            for d in range(m["dim"]):
                mesh.bc_mask_plane(mask, d, 0.0, 1e-9)
                mesh.bc_mask_plane(mask, d, 1.0, 1e-9)
        else:
            mesh.bc_mask_plane(mask, 0, 0.0, 1e-9)
        A.dirichlet(mask, b)
        x = ctx.alloc(mesh.ndofs)
        x.upload(np.zeros(mesh.ndofs))
        info = A.solve(b, x, method=hip.PCG_JACOBI, precision=hip.FP64, stop=hip.STOP_RELRES, tol=TOL)
        assert info.status == 0, info.as_dict()
    return len(mats) / (time.perf_counter() - t)


def served_stream(hip, ctx, nreq, rates, fp64_sample, n_alone, slots=128, setup_threads=0, full_at_max=None,
                  graph_rates=(50.0, 200.0, 2000.0), cg_passes=0, two_pass_limit=0):
    """c4: a served stream of Poisson / elasticity solves with Poisson
    arrivals at each rate; the record of every rate, the alone-vs-mix check,
    the plain FP64 GPU path on a sample of the same requests, and the same
    streams at graph_rates with the chunks replayed from HIP graphs."""
    from yafel_amd import workloads
    budget = 200 << 30
    nmax = max(nreq, full_at_max or 0)
    reqs = workloads.c4_requests(nmax, seed=4000)
    mats = [workloads.c4_materialise(r) for r in reqs]
    log(f"{len(mats)} requests materialised; dofs total {sum(r['ndofs'] for r in reqs)}")
    run_stream(hip, ctx, mats[:8], np.zeros(8), budget, slots, setup_threads)  # warm-up
    sweep, top = [], None
    for rate in rates:
        nr = full_at_max if (full_at_max and rate == max(rates)) else nreq
        arr = workloads.poisson_arrivals(nr, rate, seed=int(rate))
        lat, span, res, st, times = run_stream(hip, ctx, mats[:nr], arr, budget, slots, setup_threads,
                                               cg_passes=cg_passes, two_pass_limit=two_pass_limit)
        bad = sum(1 for x, inf in res.values() if inf.status != 0 or inf.rel_residual > TOL)
        rec = {"rate": rate, "requests": nr, "solves_per_s": nr / span, "p50_ms": float(np.percentile(lat, 50) * 1e3),
               "p99_ms": float(np.percentile(lat, 99) * 1e3), "failed": bad, "hbm_peak_gb": st["hbm_peak"] / 2**30,
               "batched_launches": st["batched_launches"], "two_pass_chunks": st["two_pass_chunks"],
               "three_pass_chunks": st["three_pass_chunks"], "by_class": latency_breakdown(mats[:nr], lat, times)}
        log(f"rate {rate}: { {k: v for k, v in rec.items() if k != 'by_class'} }")
        sweep.append(rec)
        if rate == max(rates):
            top = (rec, res)
    graph_sweep = []
    for rate in graph_rates:
        arr = workloads.poisson_arrivals(nreq, rate, seed=int(rate))
        lat, span, res_g, st, _ = run_stream(hip, ctx, mats[:nreq], arr, budget, slots, setup_threads, graphs=True)
        direct = next((r for r in sweep if r["rate"] == rate and r["requests"] == nreq), None)
        g = {"rate": rate, "requests": nreq, "solves_per_s": nreq / span, "p50_ms": float(np.percentile(lat, 50) * 1e3),
             "p99_ms": float(np.percentile(lat, 99) * 1e3),
             "failed": sum(1 for x, inf in res_g.values() if inf.status != 0 or inf.rel_residual > TOL),
             "graph_captures": st["graph_captures"], "graph_replays": st["graph_replays"], "chunks": st["chunks"],
             "direct_p50_ms": direct["p50_ms"] if direct else None, "direct_p99_ms": direct["p99_ms"] if direct else None}
        log(f"graph replay, rate {rate}: {g}")
        graph_sweep.append(g)
    rec, res = top
    diffs = []
    for j in range(min(n_alone, rec["requests"])):
        _, _, r1, _, _ = run_stream(hip, ctx, [mats[j]], np.zeros(1), budget, slots, setup_threads,
                                    cg_passes=cg_passes, two_pass_limit=two_pass_limit)
        diffs.append(float(np.abs(r1[0][0] - res[j][0]).max()))
    fp64_rate = fp64_one_at_a_time(hip, ctx, mats[:min(fp64_sample, rec["requests"])])
    return {"solves_per_s": rec["solves_per_s"], "p50_ms": rec["p50_ms"], "p99_ms": rec["p99_ms"],
            "rate_sweep": sweep, "graph_replay": graph_sweep,
            "alone_vs_mix_max_abs_diff": max(diffs) if diffs else None,
            "alone_vs_mix_requests": len(diffs),
            "fp64_gpu_one_at_a_time_solves_per_s": fp64_rate, "fp64_sample": min(fp64_sample, rec["requests"]),
            "speedup_vs_fp64_gpu": rec["solves_per_s"] / fp64_rate}


def serve_bench(args):
    """c4: a served stream of Poisson / elasticity solves with Poisson arrivals."""
    from yafel_amd import hip
    ctx = hip.Context(0)
    rates = [float(r) for r in args.rates.split(",")]
    graph_rates = tuple(float(r) for r in args.graph_rates.split(",") if r.strip())
    rec = served_stream(hip, ctx, args.sweep_requests, rates, args.fp64_sample, 6, args.slots, args.setup_threads,
                        full_at_max=args.requests, graph_rates=graph_rates, cg_passes=args.server_cg_passes,
                        two_pass_limit=args.server_two_pass_limit)
    nr = rec["rate_sweep"][-1]["requests"]
    out = {
        "metric": "solves/s + p99 latency at 1e-8 rel residual vs plain FP64 CG; SpMV HBM GB/s",
        "value": rec["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": nr, "warmup": 8,
        "ms_per_step": 1e3 / rec["solves_per_s"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-inner/fp64-refinement", "data": "synthetic structured meshes (seeded)",
        "config": {"workload": f"served stream of {nr} solves, Poisson arrivals at {max(rates):g}/s: 2-D Poisson Q1/Q2 "
                               "(16-256 cells/axis) or 3-D Q1 elasticity (8-48), lognormal coefficients, N(0,1) loads, "
                               "200 GB HBM budget, batched Jacobi-PCG FP32 inner + FP64 refinement to 1e-8"},
        **{k: v for k, v in rec.items() if k != "solves_per_s"},
    }
    print(json.dumps(out), flush=True)


def c3_bench(args):
    """c3: 2-D Poisson Q4 256^2 with lognormal coefficients -- the high-order
    element kernel (FP64 scalar vs FP32 matrix cores), assembly, and the solve."""
    import ctypes as C
    from yafel_amd import hip, workloads
    ctx = hip.Context(0)
    w = workloads.c3(ctx)
    nel, nen = w.mesh.nel, w.mesh.nen
    cd = ctx.upload(w.coef)
    ph_dev = hip.physics(hip.POISSON, 2, coef_dev=cd)
    K64 = ctx.alloc(nel * nen * nen)
    K32 = ctx.alloc(nel * nen * nen, np.float32)
    lib = hip.lib

    def timed(fn, reps):
        fn()
        ctx.sync()
        ctx.timer_begin()
        for _ in range(reps):
            fn()
        return ctx.timer_end() / reps

    t64 = timed(lambda: hip.check(lib.yafel_hip_element_matrices(w.mesh.h, C.byref(ph_dev), K64.ptr), ctx.h), 10)
    t32 = timed(lambda: w.mesh.element_matrices_fp32(ph_dev, K32), 20)
    issued = nel * (2 * nen // 2) * 32 * 32 * 2 * 2      # 32x32x2 steps, nen padded to 32
    useful = nel * nen * nen * 2 * nen * 2              # (2 nqp) x nen x nen, nqp = nen
    # the coefficients device-resident, like every other input of the timed region
    w.ph = ph_dev
    a64 = timed(lambda: w.assemble(), 5)
    w.extra_req = hip.FP32_ELEMENTS
    a32 = timed(lambda: w.assemble(), 5)
    # its parts: the FP32 tangent alone (element matrices written straight into
    # the matrix, shared entries summed after), and DirichletBC::apply alone
    t32k = timed(lambda: w.mesh.assemble(ph_dev, w.A, None, hip.TANGENT | hip.FP32_ELEMENTS), 10)
    tbc = timed(lambda: w.A.dirichlet(w.mask, w.b), 10)
    w.extra_req = 0
    w.assemble()
    res = {}
    for name, prec in (("fp64_pcg", hip.FP64), ("mixed_pcg", hip.MIXED)):
        info = w.solve(method=hip.PCG_JACOBI, precision=prec, stop=hip.STOP_RELRES, tol=1e-8)
        info = w.solve(method=hip.PCG_JACOBI, precision=prec, stop=hip.STOP_RELRES, tol=1e-8)
        res[name] = {"ms": info.ms, "iterations": info.iterations, "rel_residual": info.rel_residual,
                     "status": info.status}
    out = {
        "workload": "c3: 2-D Poisson Q4 (GLL) 256^2, 5x5 Gauss, k_e = exp(N(0, 0.5^2)), rhs N(0,1), "
                    "homogeneous Dirichlet on all edges",
        "element_kernel_fp64_us": t64 * 1e3, "element_kernel_fp32_mfma_us": t32 * 1e3,
        "mfma_tflops_issued": issued / (t32 * 1e-3) / 1e12, "mfma_tflops_useful": useful / (t32 * 1e-3) / 1e12,
        "mfma_peak_fp32_tflops": 157.3, "mfma_frac_issued": issued / (t32 * 1e-3) / 157.3e12,
        "fp32_staging_write_gbs": nel * nen * nen * 4 / (t32 * 1e-3) / 1e9,
        "assembly_fp64_ms": a64, "assembly_fp32_mfma_ms": a32, "fp32_tangent_ms": t32k, "dirichlet_ms": tbc,
        "solves": res,
    }
    print(json.dumps(out), flush=True)


def spmv_bench(args):
    """The SpMV alone (FP32 and FP64) on each config's matrix -- c2 BCSR3
    (the --n mesh), c3 CSR Q4, c5 CSR Q2 (with the k = 64 multi-RHS product),
    c1 CSR Q1 -- plus one copy-kernel pass of known size: the command the
    rocprof counter runs profile, the copy kernel being the calibration of
    their byte counters."""
    from yafel_amd import hip, workloads
    ctx = hip.Context(0)
    out = {"workload": "SpMV alone on the config matrices"}
    cases = (("c2", lambda: workloads.c2(ctx, n=args.n)), ("c3", lambda: workloads.c3(ctx)),
             ("c5", lambda: workloads.c5(ctx)), ("c1", lambda: workloads.c1(ctx)))
    for cname, make in cases:
        w = make()
        w.assemble()
        A = w.A
        A.sync_fp32()
        n, inf = w.n, A.info()
        ib = A.index_bytes()
        rec = {"format": "BCSR3" if inf.nblocks != inf.nnz else "CSR", "nnz": inf.nnz, "blocks": inf.nblocks,
               "index_bytes": ib}
        for name, prec, dt, s in (("fp32", hip.FP32, np.float32, 4), ("fp64", hip.FP64, np.float64, 8)):
            x = ctx.upload(np.random.default_rng(0).standard_normal(n).astype(dt))
            y = ctx.alloc(n, dt)
            A.spmv(x, y, prec)
            ctx.timer_begin()
            for _ in range(args.steps):
                A.spmv(x, y, prec)
            ms = ctx.timer_end() / args.steps
            # algorithmic bytes with int32 columns; the stored streams use ib bytes per column
            if rec["format"] == "BCSR3":
                byts, stored = bcsr_bytes(inf.nblocks, n // 3, n, s, 4), bcsr_bytes(inf.nblocks, n // 3, n, s, ib)
            else:
                byts = inf.nnz * (s + 4) + (n + 1) * 4 + 2 * n * s
                stored = inf.nnz * (s + ib) + (n + 1) * 4 + 2 * n * s
            rec[name] = {"us": ms * 1e3, "gbs": byts / (ms * 1e-3) / 1e9, "algorithmic_bytes": byts,
                         "stored_bytes": stored, "gbs_stored": stored / (ms * 1e-3) / 1e9}
        if cname == "c5":
            # the multi-RHS product on the same matrix, k = 64 right-hand sides
            k = 64
            for name, prec, dt, sz in (("spmm_fp32_k64", hip.FP32, np.float32, 4), ("spmm_fp64_k64", hip.FP64,
                                                                                      np.float64, 8)):
                X = ctx.upload(np.random.default_rng(1).standard_normal(n * k).astype(dt))
                Y = ctx.alloc(n * k, dt)
                A.spmm(X, Y, k, prec)
                ctx.timer_begin()
                for _ in range(max(1, args.steps // 5)):
                    A.spmm(X, Y, k, prec)
                ms = ctx.timer_end() / max(1, args.steps // 5)
                byts = inf.nnz * (sz + 4) + (n + 1) * 4 + 2 * k * n * sz
                rec[name] = {"us": ms * 1e3, "gbs": byts / (ms * 1e-3) / 1e9, "algorithmic_bytes": byts}
                del X, Y
        log(f"spmv {cname}: {rec}")
        out[cname] = rec
        del w, A
    out["copy_gbs"] = ctx.copy_bandwidth(1 << 30, 1)  # two launches of a 1 GiB copy
    out["read_gbs"] = ctx.read_bandwidth(1 << 30, 5)
    print(json.dumps(out), flush=True)


def c5_bench(args):
    """c5: SUPG advection-diffusion, Q2 512^2 (cell Peclet 100), 64 seeded
    N(0,1) right-hand sides (zero on the Dirichlet edges), right-Jacobi
    BiCGSTAB to 1e-8: all 64 together through the multi-RHS path (one SpMM per
    product) against one right-hand side at a time (a bounded sample of 8,
    scaled to 64)."""
    from yafel_amd import hip, workloads
    ctx = hip.Context(0)
    w = workloads.c5(ctx)
    w.assemble()
    A, n, k = w.A, w.n, w.nrhs
    A.sync_fp32()
    mask = w.mask.download().astype(bool)
    B = w.B_host.copy()
    B[mask, :] = 0.0
    Bd, Xd = ctx.upload(B.ravel()), ctx.alloc(n * k)
    out = {"workload": f"c5: SUPG Q2 512^2 ({n} dofs), {k} right-hand sides, right-Jacobi BiCGSTAB to 1e-8"}
    for name, prec in (("mixed", hip.MIXED), ("fp64", hip.FP64)):
        kw = dict(method=hip.BICGSTAB_JACOBI, precision=prec, stop=hip.STOP_RELRES, tol=TOL)
        Xd.upload(np.zeros(n * k))
        A.solve_multi(Bd, Xd, k, **kw)  # warm-up
        Xd.upload(np.zeros(n * k))
        ctx.sync()
        t0 = time.perf_counter()
        infos = A.solve_multi(Bd, Xd, k, **kw)
        ctx.sync()
        t_multi = time.perf_counter() - t0
        bad = [j for j, i in enumerate(infos) if i.status != 0 or i.rel_residual >= TOL]
        # one right-hand side at a time: 8 of the 64
        b, x = ctx.alloc(n), ctx.alloc(n)
        its, t_one = [], 0.0
        for j in range(8):
            b.upload(B[:, j])
            x.upload(np.zeros(n))
            ctx.sync()
            t0 = time.perf_counter()
            info = A.solve(b, x, **kw)
            ctx.sync()
            t_one += time.perf_counter() - t0
            its.append(info.iterations)
        out[name] = {"multi_s": t_multi, "multi_rhs_per_s": k / t_multi,
                     "multi_iterations_max": max(i.iterations for i in infos),
                     "multi_iterations_mean": float(np.mean([i.iterations for i in infos])),
                     "failed": len(bad),
                     "one_at_a_time_rhs_per_s": 8 / t_one, "one_at_a_time_iterations_mean": float(np.mean(its)),
                     "speedup_multi_vs_one": (k / t_multi) / (8 / t_one)}
        log(f"c5 {name}: {out[name]}")
    # the SpMM of the multi-RHS products alone (FP32, k = 64)
    inf = A.info()
    X32 = ctx.upload(np.random.default_rng(0).standard_normal(n * k).astype(np.float32))
    Y32 = ctx.alloc(n * k, np.float32)
    A.spmm(X32, Y32, k, hip.FP32)
    reps = 20
    ctx.timer_begin()
    for _ in range(reps):
        A.spmm(X32, Y32, k, hip.FP32)
    ms = ctx.timer_end() / reps
    byts = inf.nnz * 8 + (n + 1) * 4 + 2 * k * n * 4
    out["spmm_fp32_k64"] = {"us": ms * 1e3, "gbs": byts / (ms * 1e-3) / 1e9, "algorithmic_bytes": byts}
    print(json.dumps(out), flush=True)


def dg_bench(args):
    """The DG residual and the graph-replayed RK4 step: upwind advection
    (apps/BasicAdvection_DG, velocity (-1, 0)) on a 256 x 256 Quad4 mesh, p = 3."""
    from yafel_amd import hip
    ctx = hip.Context(0)
    n, p = 256, 3
    nodes, cells = hip.structured_mesh(2, n, n, 0, 2)  # gmsh (counter-clockwise) corners
    dg = hip.DG(ctx, p, nodes, cells)
    ph = hip.physics(hip.DG_ADVECTION, 2, U=(-1.0, 0.0, 0.0))
    X = dg.nodes()
    dx = X[:, 0] - 0.5
    U = ctx.upload(np.where(np.abs(dx) < 0.25, np.sin(4 * np.pi * dx) ** 2, 0.0))
    R = ctx.alloc(dg.ndofs)
    dg.residual(ph, U, R)
    reps = 50
    ctx.timer_begin()
    for _ in range(reps):
        dg.residual(ph, U, R)
    t_res = ctx.timer_end() / reps
    nel, nen, nf, n1 = cells.shape[0], dg.nen, dg.nfaces, p + 1
    na = dg.naffine
    # unique bytes of one evaluation: U; per cell, its 64-byte affine record and,
    # for a non-parallelogram cell, its mass inverse and node coordinates; face
    # records, node lists and geometry; the per-cell face-node flux sums written
    # and read back; the cells' face lists; R written
    byts = (dg.ndofs * 8 + nel * 64 + (nel - na) * (nen * nen * 8 + nen * 24)
            + nf * (16 + 2 * n1 * 4 + n1 * 32) + 2 * nel * 4 * n1 * 8 + nel * 16 + dg.ndofs * 8)
    dt, steps = 1e-4, 20
    dg.rk4(ph, U, dt, 1)
    ctx.sync()
    t0 = time.perf_counter()
    dg.rk4(ph, U, dt, steps)
    t_rk = (time.perf_counter() - t0) / steps * 1e3
    out = {"workload": f"DG upwind advection, Quad4 {n}^2, p = {p} ({dg.ndofs} dofs, {nf} faces)",
           "affine_cells": na,
           "residual_us": t_res * 1e3, "residual_gbs": byts / (t_res * 1e-3) / 1e9,
           "residual_bytes": byts, "elements_per_s": nel / (t_res * 1e-3),
           "rk4_step_ms": t_rk, "rk4_steps_per_s": 1e3 / t_rk}
    print(json.dumps(out), flush=True)


def recovery_bench(args):
    """Gradient recovery after a solve, at the sizes of the apps that call it:
    LocalSmoothingGradient on 3-D elasticity Q1 64^3 (apps/BasicLinearElasticity,
    823,875 dofs) and ZZGradientRecovery on 2-D Poisson Q1 1024^2
    (apps/BasicPoisson, 1,050,625 nodes), seeded N(0,1) fields.  The oracle's
    serial restatement of each is timed on the same inputs beside them."""
    from yafel_amd import hip
    ctx = hip.Context(0)
    out = {"workload": "gradient recovery: local smoothing, elasticity Q1 64^3 (dpn 3); ZZ, Poisson Q1 1024^2"}
    reps = 20
    cases = (("local_smoothing", 3, 64, 3, "gradient_local_smoothing"), ("zz", 2, 1024, 1, "gradient_zz"))
    for name, dim, n, dpn, fn in cases:
        m = hip.Mesh.structured(ctx, dim, 1, dpn, n)
        u = ctx.upload(hip.normal(7000 + dim, m.ndofs))
        rows = m.ndofs if name == "local_smoothing" else m.n_nodes
        G = ctx.alloc(rows * dim)
        getattr(m, fn)(u, G)
        ctx.timer_begin()
        for _ in range(reps):
            getattr(m, fn)(u, G)
        ms = ctx.timer_end() / reps
        # unique bytes: connectivity, coordinates, the field, the gradient written
        uniq = m.nel * m.nen * 4 + m.n_nodes * 24 + m.ndofs * 8 + rows * dim * 8
        # the staged per-element values written once and read once
        stage = (m.nel * m.nen * (1 + dpn * dim) if name == "local_smoothing" else m.nel * m.nen * 2 * dim) * 8
        rec = {"dofs": m.ndofs, "elements": m.nel, "ms": ms, "unique_bytes": uniq, "staged_bytes": stage,
               "gbs_unique": uniq / (ms * 1e-3) / 1e9, "gbs_with_staging": (uniq + 2 * stage) / (ms * 1e-3) / 1e9}
        if not args.no_cpu:
            from oracle.oracle import Oracle
            orc = Oracle(fast=True)
            el, dn = m.download()
            uh = u.download()
            t0 = time.perf_counter()
            if name == "local_smoothing":
                orc.grad_local_smoothing(dim, 1, dpn, dn, el, uh)
            else:
                orc.grad_zz(dim, dn, el, uh)
            t_cpu = time.perf_counter() - t0
            rec["cpu_baseline"] = {"ms": t_cpu * 1e3, "kind": "port", "cores": 1,
                                   "sample": "the oracle's serial restatement on the same mesh and field"}
        log(f"{name}: {rec}")
        out[name] = rec
    print(json.dumps(out), flush=True)


def multicg_bench(args):
    """Multi-RHS Jacobi-PCG on 3-D elasticity Q1 32^3 in CSR (107,811 dofs),
    seeded N(0,1) right-hand sides zero on the clamped face, to 1e-8: MIXED
    with k = 64 (four right-hand sides per lane) and k = 63 (one per lane),
    FP64 with k = 64 -- the batches behind the multi-RHS CG product figures."""
    from yafel_amd import hip, workloads
    ctx = hip.Context(0)
    w = workloads.c2(ctx, n=32, fmt=hip.CSR)
    w.assemble()
    A, n = w.A, w.n
    A.sync_fp32()
    mask = w.mask.download().astype(bool)
    out = {"workload": f"multi-RHS Jacobi-PCG, elasticity Q1 32^3 CSR ({n} dofs), to 1e-8"}
    for name, prec, k in (("mixed_k64", hip.MIXED, 64), ("mixed_k63", hip.MIXED, 63), ("fp64_k64", hip.FP64, 64)):
        B = np.stack([hip.normal(7100 + j, n) for j in range(k)], axis=1)
        B[mask, :] = 0.0
        Bd, Xd = ctx.upload(np.ascontiguousarray(B).ravel()), ctx.alloc(n * k)
        kw = dict(method=hip.PCG_JACOBI, precision=prec, stop=hip.STOP_RELRES, tol=TOL)
        Xd.upload(np.zeros(n * k))
        A.solve_multi(Bd, Xd, k, **kw)
        Xd.upload(np.zeros(n * k))
        ctx.sync()
        t0 = time.perf_counter()
        infos = A.solve_multi(Bd, Xd, k, **kw)
        ctx.sync()
        dt = time.perf_counter() - t0
        out[name] = {"ms": dt * 1e3, "iterations_mean": float(np.mean([i.iterations for i in infos])),
                     "failed": sum(1 for i in infos if i.status != 0 or i.rel_residual >= TOL)}
        log(f"{name}: {out[name]}")
    print(json.dumps(out), flush=True)


def c2_bench(args):
    """The default line: c2 solves/s with the FP64 GPU paths, the roofline of
    the CG SpMV, a bounded c4 served stream and the reference CPU path."""
    # counters first, in child processes, before this process touches the GPU
    traffic, traffic_detail = (None, {"error": "skipped (--no-traffic)"})
    if not args.no_traffic:
        traffic, traffic_detail = measured_traffic(args.n)
    from yafel_amd import hip, workloads
    ctx = hip.Context(0)
    t = time.perf_counter()
    w = workloads.c2(ctx, n=args.n)
    setup_s = time.perf_counter() - t
    log(f"setup {setup_s:.2f}s  dofs={w.n}")
    mixed = dict(method=hip.PCG_JACOBI, precision=hip.MIXED, stop=hip.STOP_RELRES, tol=TOL)

    for i in range(args.warmup):
        info = w.step(**mixed)
        log(f"warmup {i}: {info.as_dict()}")
    ctx.sync()
    lat, infos = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        info = w.step(**mixed)
        lat.append(time.perf_counter() - ts)
        infos.append(info)
    ctx.sync()
    T = time.perf_counter() - t0
    bad = [i.as_dict() for i in infos if i.status != 0 or i.rel_residual > TOL]
    if bad:
        log(f"NOT CONVERGED: {bad[:2]}")
        sys.exit(1)
    log(f"timed {args.steps} steps in {T:.3f}s; last {infos[-1].as_dict()}")
    xs = {"mixed_pcg": w.x.download()}

    # plain FP64 GPU paths (the correctness baseline): same step with FP64 arithmetic
    fp64 = {}
    if not args.no_fp64:
        for name, meth in [("cg", hip.CG), ("pcg_jacobi", hip.PCG_JACOBI)]:
            ctx.sync()
            ts = time.perf_counter()
            info = w.step(method=meth, precision=hip.FP64, stop=hip.STOP_RELRES, tol=TOL)
            dt = time.perf_counter() - ts
            fp64[name] = {"solves_per_s": 1.0 / dt, "ms": dt * 1e3, "iterations": info.iterations,
                          "rel_residual": info.rel_residual, "status": hip.STATUS[info.status]}
            xs["fp64_" + name] = w.x.download()
            log(f"fp64 {name}: {fp64[name]}")
        ctx.sync()

    # roofline of the dominant kernel: the CG iteration's FP32 SpMV (q = A p
    # and the p.q partials), every launch of one more step timed with its own
    # HIP event pair on the context stream; launches that returned at once
    # (after convergence or at a refinement pause) are left out
    A = w.A
    inf = A.info()
    nbr = inf.rows // 3
    ctx.kernel_timing(True)
    w.step(**mixed)
    ka_ms, ka_n = ctx.kernel_time()
    ctx.kernel_timing(False)
    ka_us = ka_ms / max(ka_n, 1) * 1e3
    alg_bytes = bcsr_bytes(inf.nblocks, nbr, w.n, 4, 4)  # matrix with int32 block columns, p read, q written
    own_bytes = bcsr_bytes(inf.nblocks, nbr, w.n, 4, A.index_bytes())  # as stored (16-bit column offsets)
    ka_gbs = alg_bytes / (ka_us * 1e-6) / 1e9
    log(f"kernel A: {ka_n} launches that ran the product, {ka_us:.2f} us each, {ka_gbs:.0f} GB/s")
    x32 = ctx.upload(np.random.default_rng(0).standard_normal(w.n).astype(np.float32))
    y32 = ctx.alloc(w.n, np.float32)
    A.sync_fp32()
    reps = 200
    for _ in range(10):
        A.spmv(x32, y32, hip.FP32)
    ctx.timer_begin()
    for _ in range(reps):
        A.spmv(x32, y32, hip.FP32)
    ms32 = ctx.timer_end() / reps
    x64 = ctx.upload(np.random.default_rng(0).standard_normal(w.n))
    y64 = ctx.alloc(w.n)
    ctx.timer_begin()
    for _ in range(reps):
        A.spmv(x64, y64, hip.FP64)
    ms64 = ctx.timer_end() / reps
    gbs32 = bcsr_bytes(inf.nblocks, nbr, w.n, 4, 4) / (ms32 * 1e-3) / 1e9
    gbs64 = bcsr_bytes(inf.nblocks, nbr, w.n, 8, 4) / (ms64 * 1e-3) / 1e9
    copy_gbs = ctx.copy_bandwidth(1 << 30, 20)
    read_gbs = ctx.read_bandwidth(1 << 30, 20)
    log(f"spmv fp32 {ms32 * 1e3:.1f}us {gbs32:.0f} GB/s; fp64 {ms64 * 1e3:.1f}us {gbs64:.0f} GB/s; "
        f"copy {copy_gbs:.0f} GB/s; read {read_gbs:.0f} GB/s")
    del x32, y32, x64, y64

    served = None
    if not args.no_serve:
        rates = [float(r) for r in args.serve_rates.split(",")]
        served = served_stream(hip, ctx, args.serve_requests, rates, args.serve_fp64_sample, 4, args.slots,
                               args.setup_threads)
        log(f"served: { {k: v for k, v in served.items() if k != 'rate_sweep'} }")

    cpu = None
    if not args.no_cpu:
        iters = fp64.get("cg", {}).get("iterations") or infos[-1].iterations
        cpu = cpu_baseline(w, xs, iters)
        log(f"cpu baseline: {cpu}")

    lat_ms = np.array(lat) * 1e3
    value = args.steps / T
    out = {
        "metric": "solves/s + p99 latency at 1e-8 rel residual vs plain FP64 CG; SpMV HBM GB/s",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": T / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32-inner/fp64-refinement",
        "data": "synthetic structured mesh (deterministic)",
        "config": {"workload": f"3-D linear elasticity Q1 hex {args.n}^3, BCSR3, Jacobi-PCG "
                               f"FP32 inner + FP64 refinement to 1e-8 (assembly + Dirichlet + solve per step)",
                   "dofs": w.n, "blocks": inf.nblocks, "tol": TOL},
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "iterations": infos[-1].iterations, "refinements": infos[-1].refinements,
        "rel_residual": infos[-1].rel_residual, "setup_s": setup_s,
        "fp64_gpu": fp64,
        "speedup_vs_fp64_cg": (value / fp64["cg"]["solves_per_s"]) if "cg" in fp64 else None,
        "spmv_fp32_gbs": gbs32, "spmv_fp32_us": ms32 * 1e3, "spmv_fp64_gbs": gbs64, "spmv_fp64_us": ms64 * 1e3,
        "copy_gbs": copy_gbs, "read_gbs": read_gbs,
        "roofline": {"bound": "hbm", "achieved": ka_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": ka_gbs / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": KERNEL_A, "bytes_per_launch": alg_bytes, "stored_bytes_per_launch": own_bytes,
                     "us_per_launch": ka_us, "launches_timed": ka_n, "traffic_detail": traffic_detail},
        "served_stream": served,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="cells per axis (64 = the benchmark config)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter runs")
    ap.add_argument("--no-serve", action="store_true", help="skip the bounded served stream")
    ap.add_argument("--workload", default="c2",
                    choices=["c2", "serve", "c3", "spmv", "dg", "c5", "recovery", "multicg", "traffic-probe"])
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--sweep-requests", type=int, default=300)
    ap.add_argument("--rates", default="50,200,500,1000,2000")
    ap.add_argument("--serve-requests", type=int, default=300, help="default line: requests per rate")
    ap.add_argument("--serve-rates", default="50,200,2000", help="default line: arrival rates")
    ap.add_argument("--serve-fp64-sample", type=int, default=40, help="default line: FP64 one-at-a-time sample")
    ap.add_argument("--graph-rates", default="50,200,2000",
                    help="serve: rates also run with chunks replayed from HIP graphs ('' for none)")
    ap.add_argument("--server-cg-passes", type=int, default=0,
                    help="serve: kernels per batched CG iteration, 2 or 3 (0: two while the batch is small)")
    ap.add_argument("--server-two-pass-limit", type=int, default=0,
                    help="serve: stored scalars of the active slots up to which a chunk takes two passes (0: default)")
    ap.add_argument("--slots", type=int, default=128)
    ap.add_argument("--fp64-sample", type=int, default=100)
    ap.add_argument("--setup-threads", type=int, default=0, help="server setup workers (0: library default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): each solve of this workload "
              "fits one MI355X, and multi-GPU sharding is not part of this build", file=sys.stderr)
        sys.exit(2)
    modes = {"serve": serve_bench, "c3": c3_bench, "spmv": spmv_bench, "c5": c5_bench, "recovery": recovery_bench,
             "dg": dg_bench, "multicg": multicg_bench, "traffic-probe": traffic_probe, "c2": c2_bench}
    modes[args.workload](args)


if __name__ == "__main__":
    main()
