"""Benchmark of the LPM solve path on one MI355X.

Workload (one step = one solve): 256^3 simple-cubic LPM (16.8 M particles, 50 M DOF), four-phase cement-like
microstructure of random overlapping spheres (porous hydrate Kn 0.1 / hydrate 1 / portlandite 2 / clinker 5 at
0.15 / 0.55 / 0.15 / 0.15, Tv = 0.3 Kn, two-end mean per bond), tension along z (bottom layer clamped, top layer
translated), block-Jacobi PCG with FP32 inner iterations inside FP64 iterative refinement to a true FP64 relative
residual of 1e-8.  A step = right-hand side + preconditioner + refined PCG, inputs resident in HBM.

Prints one JSON line: solves/s, latency percentiles, the K u kernel's roofline (HIP events around every K u launch
of the timed solves; algorithmic bytes N (2 d w + 18 w) per launch) and the CPU baseline (the C restatement of the
reference's block-Jacobi PCG, FP64, on the host cores, on a bounded number of iterations of the same problem).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lpm-solver-trained-coarse-graining_amd"))

METRIC = "K·u HBM GB/s (% peak); CG solves/s + p99 latency at rel-res 1e-8"
HBM_PEAK_GBS = 8000.0
PHASE_KN = np.array([0.1, 1.0, 2.0, 5.0])
# HBM bytes per launch of the masked K p kernel at 256^3 (FP32, packed), from rocprofv3 --pmc FETCH_SIZE and
# --pmc WRITE_SIZE runs of `bench.py --probe 256` (KiB per dispatch), scaled by the same counters of a 1.61 GB copy
# (factors 2.0 and 1.0; profiles/ku3d_256cube_fp32_packed.md)
KU_TRAFFIC_256 = 802852 * 1024 * 2.0 + 196611 * 1024 * 1.0


def probe(n, reps):
    """Kernel probe for the profilers: reps x (calibration copy of 1.61 GB, plain K u), then reps PCG iterations
    (masked K p inside the solver) on an n^3 lattice with random symmetric coefficients."""
    import torch
    import synth
    import lpm_ops as ops
    import _lpmlib as L
    dev = torch.device("cuda:0")
    dims = (n, n, n)
    N = n ** 3
    pp = np.random.default_rng(0).uniform(0.5, 2.0, (N, 3))
    kp, tp = synth.packed_from_particles(dims, pp, dev, torch.float32)
    u = torch.randn(N, 3, device=dev)
    free = torch.ones(N, device=dev)
    x = torch.ones(N * 24, device=dev)
    y = torch.empty_like(x)
    minv = ops.block_jacobi(dims, 1.0, kp, tp, free=free, layout=L.PACKED)
    for _ in range(reps):
        torch.mul(x, 2.0, out=y)
        ops.apply(dims, 1.0, u, kp, tp, layout=L.PACKED)
    ops.pcg(dims, 1.0, torch.randn(N, 3, device=dev), kp, tp, minv=minv, free=free, tol=0.0, maxit=reps,
            layout=L.PACKED)
    torch.cuda.synchronize()
    print("[probe] done", file=sys.stderr, flush=True)


def build_problem(n, seed, dev):
    import torch
    import synth
    import lpm_ops as ops
    import _lpmlib as L
    dims = (n, n, n)
    t0 = time.perf_counter()
    lab = synth.cement_phases(n, seed)
    kn = PHASE_KN[lab]
    ppar = np.stack([kn, kn, 0.3 * kn], 1)
    kp64, tp64 = synth.packed_from_particles(dims, ppar, dev, torch.float64)
    kp32, tp32 = kp64.float(), tp64.float()
    z = torch.arange(n, device=dev).repeat_interleave(n * n)
    fixed = (z == 0) | (z == n - 1)
    up = torch.zeros((n ** 3, 3), device=dev, dtype=torch.float64)
    up[z == n - 1, 2] = 1e-3 * (n - 1)
    free = (~fixed).to(torch.float64)
    t_setup = time.perf_counter() - t0
    return dict(dims=dims, lab=lab, kp64=kp64, tp64=tp64, kp32=kp32, tp32=tp32, up=up, free=free, fixed=fixed,
                t_setup=t_setup, ops=ops, L=L, fractions=np.bincount(lab, minlength=4) / lab.size)


def solve(P, tol):
    import torch
    ops, L = P["ops"], P["L"]
    dims = P["dims"]
    b = -ops.apply(dims, 1.0, P["up"], P["kp64"], P["tp64"], layout=L.PACKED) * P["free"][:, None]
    M = ops.block_jacobi(dims, 1.0, P["kp64"], P["tp64"], free=P["free"], out_dtype=torch.float32, layout=L.PACKED)
    x, info = ops.pcg_refined(dims, 1.0, b, P["kp64"], P["tp64"], P["kp32"], P["tp32"], minv32=M, free=P["free"],
                              tol=tol, maxit=100000, inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED)
    return x, info


def _cpu_problem(P):
    """Host copies of the 256^3 problem for the C oracle: (lib, nbr, Kn, Tv, M (N,3,3), free (N,) uint8, b)."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build_oracle
    import synth
    import lpm_ops as ops
    if not os.path.exists(build_oracle.LIB):
        build_oracle.build()
    lib = ctypes.CDLL(build_oracle.LIB)
    lib.orc_threads.restype = ctypes.c_int
    dims = P["dims"]
    n = dims[0]
    N = n ** 3
    Kn, Tv = synth.directed_from_packed(dims, P["kp64"], P["tp64"])
    Kn = np.ascontiguousarray(Kn.cpu().numpy())
    Tv = np.ascontiguousarray(Tv.cpu().numpy())
    M = ops.full_from_sym(ops.block_jacobi(dims, 1.0, P["kp64"], P["tp64"], free=P["free"], layout=P["L"].PACKED))
    M = np.ascontiguousarray(M.cpu().numpy())
    b = -ops.apply(dims, 1.0, P["up"], P["kp64"], P["tp64"], layout=P["L"].PACKED) * P["free"][:, None]
    b = np.ascontiguousarray(b.cpu().numpy())
    fr = np.ascontiguousarray((~P["fixed"]).cpu().numpy().astype(np.uint8))
    del torch
    p = np.arange(N, dtype=np.int64)
    ijk = np.stack([p % n, (p // n) % n, p // (n * n)], 1)
    nbr = np.empty((N, 18), np.int64)
    offs = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
                     (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1),
                     (0, 1, -1), (0, -1, 1)])
    for k, o in enumerate(offs):
        q = ijk + o
        ok = np.all((q >= 0) & (q < n), 1)
        nbr[:, k] = np.where(ok, (q[:, 2] * n + q[:, 1]) * n + q[:, 0], -1)
    return lib, nbr, Kn, Tv, M, fr, b


def amg_cpu_baseline(P, gpu_iters, sample_its=3):
    """The C/OpenMP restatement of the reference's AMG-PCG (oracle/lpm_oracle_amg.c, FP64) on the same problem:
    the whole numeric setup plus `sample_its` PCG iterations are timed; the solve time is the setup plus the time
    per iteration times the FP64 GPU AMG-PCG's iteration count.  The block-Jacobi inverse of the finest level comes
    from the GPU (as in the block-Jacobi baseline) and is not timed."""
    lib, nbr, Kn, Tv, M, fr, b = _cpu_problem(P)
    import amg_ref                                  # oracle/, on the path after _cpu_problem
    dims = P["dims"]
    n = dims[0]
    lo, L_ = (0, 0, 1), (n, n, n - 2)
    t0 = time.perf_counter()
    c = amg_ref.CAmg(lib, dims, lo, L_, nbr, Kn, Tv, M, fr.astype(bool), **{k: v for k, v in
                                                                             {"agg": 3, "coarse_max": 1000, "nu": 1,
                                                                              "gamma": 2, "power_its": 5}.items()})
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, done, _ = c.pcg(b, 0.0, sample_its)
    t_it = (time.perf_counter() - t0) / max(done, 1)
    t_solve = t_setup + t_it * gpu_iters
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": int(lib.orc_threads()), "kind": "port",
            "sample": "numeric setup (%.1f s: rho by power iteration, level-1 stencil by probing with the C K u, "
                      "coarse Galerkin products, inverses) and %d W-cycle AMG-PCG iterations (%.3f s each) of the C/"
                      "OpenMP restatement of timing/t3d/amg3d.py:508-672 on the same 256^3 problem (FP64), scaled "
                      "by the FP64 GPU AMG-PCG's %d iterations" % (t_setup, done, t_it, gpu_iters),
            "setup_s": t_setup, "sec_per_iter": t_it}


def host_cpu():
    """(model name, logical CPUs this process may use, threads for 'all cores': the affinity set, capped by
    OMP_NUM_THREADS when the host sets it)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return model, avail, min(avail, cap) if cap > 0 else avail


def fp64_pcg_iterations(P, tol):
    """Iterations of the FP64 block-Jacobi PCG (the reference's algorithm and arithmetic, coarse3d/lin3d.py:464-518)
    on the same problem, run once on the GPU outside the timed region: the count the CPU baseline is scaled by."""
    import torch
    ops, L = P["ops"], P["L"]
    dims = P["dims"]
    b = -ops.apply(dims, 1.0, P["up"], P["kp64"], P["tp64"], layout=L.PACKED) * P["free"][:, None]
    M = ops.block_jacobi(dims, 1.0, P["kp64"], P["tp64"], free=P["free"], layout=L.PACKED)
    _, info = ops.pcg(dims, 1.0, b, P["kp64"], P["tp64"], minv=M, free=P["free"], tol=tol, maxit=100000,
                      layout=L.PACKED)
    torch.cuda.synchronize()
    return info.it, info.relmax


def cpu_baseline(P, gpu_iters, budget_s=20.0, tol=1e-8, rounds=3):
    """The C oracle's block-Jacobi PCG (FP64, OpenMP: the reference's CPU solver, coarse3d/lin3d.py:464-518) on the
    same problem: time per iteration as the median of `rounds` rounds of a bounded number of iterations (warm-up
    excluded, timing/summ_timing.py:19-46), at 6 threads (the reference's LPMCG_THREADS, timing/t2d/a2common.py:10)
    and at every core this process may use; the solve time is that times the iteration count of an FP64 PCG solve
    of the same problem (run on the GPU; the CPU solver takes the same iterations in the same arithmetic)."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build_oracle
    import synth
    if not os.path.exists(build_oracle.LIB):
        build_oracle.build()
    lib = ctypes.CDLL(build_oracle.LIB)
    Pp = ctypes.c_void_p
    lib.orc_pcg3.argtypes = [ctypes.c_int64, Pp, Pp, Pp, Pp, Pp, Pp, Pp, ctypes.c_double, ctypes.c_int, Pp]
    lib.orc_pcg3.restype = ctypes.c_int
    lib.orc_threads.restype = ctypes.c_int
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    dims = P["dims"]
    n = dims[0]
    N = n ** 3
    it64, rel64 = fp64_pcg_iterations(P, tol)
    Kn, Tv = synth.directed_from_packed(dims, P["kp64"], P["tp64"])
    Kn = np.ascontiguousarray(Kn.cpu().numpy())
    Tv = np.ascontiguousarray(Tv.cpu().numpy())
    import lpm_ops as ops
    M = ops.full_from_sym(ops.block_jacobi(dims, 1.0, P["kp64"], P["tp64"], free=P["free"], layout=P["L"].PACKED))
    M = np.ascontiguousarray(M.cpu().numpy())
    b = -ops.apply(dims, 1.0, P["up"], P["kp64"], P["tp64"], layout=P["L"].PACKED) * P["free"][:, None]
    b = np.ascontiguousarray(b.cpu().numpy())
    fr = np.ascontiguousarray((~P["fixed"]).cpu().numpy().astype(np.uint8))
    del torch
    # neighbour table (N, 18) int64, -1 outside the box (the reference's lattice.nbr)
    p = np.arange(N, dtype=np.int64)
    ijk = np.stack([p % n, (p // n) % n, p // (n * n)], 1)
    nbr = np.empty((N, 18), np.int64)
    offs = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
                     (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1),
                     (0, 1, -1), (0, -1, 1)])
    for k, o in enumerate(offs):
        q = ijk + o
        ok = np.all((q >= 0) & (q < n), 1)
        nbr[:, k] = np.where(ok, (q[:, 2] * n + q[:, 1]) * n + q[:, 0], -1)
    del ijk, p
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rel = np.zeros(1)
    x = np.zeros_like(b)
    model, avail, all_threads = host_cpu()
    per = {}
    for nthr in sorted({min(6, all_threads), all_threads}):
        lib.orc_set_threads(nthr)
        x[:] = 0
        t0 = time.perf_counter()
        lib.orc_pcg3(N, ptr(nbr), ptr(Kn), ptr(Tv), ptr(M), ptr(fr), ptr(b), ptr(x), 0.0, 2, ptr(rel))   # warm-up
        t_probe = (time.perf_counter() - t0) / 2
        # about budget_s / 2 of CPU work per thread count, split over the rounds
        its = int(max(3, min(200, budget_s / 2 / rounds / max(t_probe, 1e-6))))
        times = []
        for _ in range(rounds):
            x[:] = 0
            t0 = time.perf_counter()
            done = lib.orc_pcg3(N, ptr(nbr), ptr(Kn), ptr(Tv), ptr(M), ptr(fr), ptr(b), ptr(x), 0.0, its, ptr(rel))
            times.append((time.perf_counter() - t0) / max(done, 1))
        per[nthr] = dict(sec_per_iter=float(np.median(times)), rounds=[float(v) for v in times], iters_per_round=its,
                         solves_per_s=1.0 / (float(np.median(times)) * it64))
    t_it = per[all_threads]["sec_per_iter"]
    out = {"value": 1.0 / (t_it * it64), "unit": "solves/s", "cores": all_threads, "kind": "port",
           "sample": "block-Jacobi PCG (FP64, C/OpenMP restatement of coarse3d/lin3d.py:464-518) on the same %d^3 "
                     "problem: %d iterations per round, median of %d rounds, %.3f s per iteration at %d threads "
                     "(%s); solve time = that x %d iterations of an FP64 block-Jacobi PCG solve of the problem to "
                     "rel-res %g (the GPU's mixed-precision solve ran %d inner iterations)"
                     % (n, per[all_threads]["iters_per_round"], rounds, t_it, all_threads, model, it64, tol,
                        gpu_iters),
           "cpu_model": model, "threads_available": avail, "fp64_pcg_iters": it64, "fp64_pcg_relres": rel64,
           "sec_per_iter": t_it, "per_threads": {str(k): v for k, v in per.items()}}
    out["fp64_oracle_64cube"] = oracle_subvolume_check(lib, P["lab"], n)
    return out


def oracle_subvolume_check(lib, lab, n_full, n=64):
    """Error of the GPU solve path against the FP64 oracle on the n^3 corner sub-volume of the same microstructure
    (tension z): FP32 PCG + FP64 refinement to 1e-8 on the GPU vs the C oracle's FP64 block-Jacobi PCG to 1e-13
    (a sparse direct factorisation of 786k DOF is out of budget)."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lpm_ref as R
    import synth
    import lpm_ops as ops
    import _lpmlib as L
    dev = torch.device("cuda:0")
    sub = lab.reshape(n_full, n_full, n_full)[:n, :n, :n].ravel()
    kn = PHASE_KN[sub]
    ppar = np.stack([kn, kn, 0.3 * kn], 1)
    dims = (n, n, n)
    lat = R.Box(dims)
    Kn, Tv = R.two_end_mean(lat, ppar)
    z = lat.ijk[:, 2]
    fixed = (z == 0) | (z == n - 1)
    up = np.zeros((lat.N, 3))
    up[z == n - 1, 2] = 1e-3 * (n - 1)
    kp64, tp64 = synth.packed_from_particles(dims, ppar, dev, torch.float64)
    upt = torch.as_tensor(up, device=dev)
    free = torch.as_tensor((~fixed).astype(float), device=dev)
    b = -ops.apply(dims, 1.0, upt, kp64, tp64, layout=L.PACKED) * free[:, None]
    M = ops.block_jacobi(dims, 1.0, kp64, tp64, free=free, out_dtype=torch.float32, layout=L.PACKED)
    x, info = ops.pcg_refined(dims, 1.0, b, kp64, tp64, kp64.float(), tp64.float(), minv32=M, free=free, tol=1e-8,
                              maxit=100000, inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED)
    ug = up + x.cpu().numpy()
    bo = np.ascontiguousarray(-R.apply(lat, up, Kn, Tv) * (~fixed)[:, None])
    Mo = np.ascontiguousarray(R.precond(lat, Kn, Tv, ~fixed))
    nbr = np.ascontiguousarray(lat.nbr.astype(np.int64))
    fr = np.ascontiguousarray((~fixed).astype(np.uint8))
    xo = np.zeros_like(bo)
    rel = np.zeros(1)
    pt = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    it = lib.orc_pcg3(lat.N, pt(nbr), pt(np.ascontiguousarray(Kn)), pt(np.ascontiguousarray(Tv)), pt(Mo), pt(fr),
                      pt(bo), pt(xo), 1e-13, 100000, pt(rel))
    uo = up + xo
    so, sg = R.stress(lat, uo, Kn, Tv), R.stress(lat, ug, Kn, Tv)
    top = z == n - 1
    Fo, Fg = R.apply(lat, uo, Kn, Tv)[top, 2].sum(), R.apply(lat, ug, Kn, Tv)[top, 2].sum()
    return {"subvolume": "%d^3 corner, tension z" % n, "gpu_relres_fp64": float(info.relmax),
            "gpu_inner_iters": info.it, "oracle_iters": int(it), "oracle_relres": float(rel[0]),
            "rel_err_u": float(np.abs(ug - uo).max() / np.abs(uo).max()),
            "rel_err_stress": float(np.abs(sg - so).max() / np.abs(so).max()),
            "rel_err_reaction": float(abs(Fg - Fo) / abs(Fo))}


# ------------------------------------------------------------------------------------------------ other workloads
def _oracle_lib():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build_oracle
    if not os.path.exists(build_oracle.LIB):
        build_oracle.build()
    lib = ctypes.CDLL(build_oracle.LIB)
    Pp = ctypes.c_void_p
    lib.orc_cg2.argtypes = [ctypes.c_int64, Pp, Pp, Pp, Pp, Pp, Pp, ctypes.c_double, ctypes.c_int, Pp]
    lib.orc_cg2.restype = ctypes.c_int
    lib.orc_pcg3.argtypes = [ctypes.c_int64, Pp, Pp, Pp, Pp, Pp, Pp, Pp, ctypes.c_double, ctypes.c_int, Pp]
    lib.orc_pcg3.restype = ctypes.c_int
    lib.orc_threads.restype = ctypes.c_int
    return lib


def _line(metric, value, unit, steps, warmup, ms, dtype, workload, extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": {"workload": workload}}
    out.update(extra)
    print(json.dumps(out), flush=True)


def bench_plain2d(a):
    """128^2 plain 2D LPM, FP64 unpreconditioned CG to 1e-8: device-driven batched solver, the host-driven plain
    GPU CG, and the C restatement of cg2plain.cg_mf on the host cores."""
    import torch
    import lpm_ops as ops
    import _lpmlib as L
    import serve
    import workloads
    dev = torch.device("cuda:0")
    P = workloads.plain2d(128, a.seed, dev)
    dims = P["dims"]
    kp, tp, _ = ops.pack(dims, P["Kn"], P["Tv"])
    mfree = P["mask"]
    b = -ops.apply(dims, 1.0, P["up"], P["Kn"], P["Tv"]) * mfree.double()

    # the single-launch persistent CG (one cooperative launch per solve) is the timed path
    cgp = ops.Cg2Persistent(dims, P["Kn"], P["Tv"], mfree)

    def gpu_solve():
        return cgp.solve(b, tol=a.tol, maxit=200000)

    def batched_solve():
        return ops.pcg(dims, 1.0, b, kp, tp, minv=None, free=mfree, tol=a.tol, maxit=200000, check=50,
                       graph=True, layout=L.PACKED)
    for _ in range(a.warmup):
        gpu_solve()
    torch.cuda.synchronize()
    lat = []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        x, info = gpu_solve()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    x = x.reshape(-1, 2)
    xb, info_b = batched_solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xb, info_b = batched_solve()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    # the pipelined form (one reduction per iteration, hidden behind the product): opt-in, its own accuracy
    cpp = ops.Cg2Persistent(dims, P["Kn"], P["Tv"], mfree, pipelined=True)
    xp, info_p = cpp.solve(b, tol=a.tol, maxit=200000)
    lat_p = []
    for _ in range(max(3, a.steps)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xp, info_p = cpp.solve(b, tol=a.tol, maxit=200000)
        torch.cuda.synchronize()
        lat_p.append(time.perf_counter() - t0)
    diff_p = float((xp.reshape(-1, 2) - x).abs().max() / x.abs().max())
    diff_b = float((xb.reshape(-1, 2) - x).abs().max() / x.abs().max())
    req = serve.Request(128, 128, 1.0, P["Kn"], P["Tv"], mfree, P["up"], "tension")
    t0 = time.perf_counter()
    up_plain, it_plain = serve.plain_cg(req, a.tol)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    diff = float((up_plain - (P["up"] + x)).abs().max() / (P["up"] + x).abs().max())
    lib = _oracle_lib()
    lat2 = P["lat"]
    nbr = np.ascontiguousarray(lat2.nbr.astype(np.int64))
    bb = np.ascontiguousarray(b.cpu().numpy().ravel())
    xx = np.zeros_like(bb)
    m = np.ascontiguousarray(P["mask_np"].astype(np.float64).ravel())
    rel = np.zeros(1)
    pt = lambda v: v.ctypes.data_as(ctypes.c_void_p)
    t0 = time.perf_counter()
    it_cpu = lib.orc_cg2(lat2.N, pt(nbr), pt(np.ascontiguousarray(P["Kn_np"])), pt(np.ascontiguousarray(P["Tv_np"])),
                         pt(m), pt(bb), pt(xx), a.tol, 200000, pt(rel))
    t_cpu = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1e3
    _line(METRIC, a.steps / sum(lat), "solves/s", a.steps, a.warmup, float(lat_ms.mean()), "f64",
          "plain 2D LPM 128^2 (32768 DOF), disks phi 0.30 contrast 10, unpreconditioned CG to rel-res %g" % a.tol,
          {"latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
           "solve": {"cg_iters": info.it, "relres": float(info.relmax), "phase_fraction": P["phase_fraction"],
                     "path": "lpm_cg2_persistent (one cooperative launch, 64 workgroups of 2 rows, 512 threads)"},
           "pipelined_cg": {"solves_per_s": len(lat_p) / sum(lat_p), "iters": info_p.it,
                            "relres_true": float(info_p.relmax), "max_rel_diff": diff_p,
                            "note": "lpm_cg2_pipelined (opt-in): one grid-wide reduction per iteration overlapped with "
                                    "the product, true residual confirmed; its iterates drift from CG's"},
           "batched_graph_cg": {"solves_per_s": 1.0 / t_batched, "iters": info_b.it, "max_rel_diff": diff_b,
                                "note": "lpm_pcg_solve, launch per kernel step replayed from a HIP graph"},
           "plain_gpu_cg": {"solves_per_s": 1.0 / t_plain, "iters": it_plain, "max_rel_diff": diff,
                            "note": "host-driven alpha/beta, one solve, FP64, same HIP K u"},
           "cpu_baseline": {"value": 1.0 / t_cpu, "unit": "solves/s", "cores": int(lib.orc_threads()), "kind": "port",
                            "sample": "full solve, %d CG iterations (C/OpenMP restatement of timing/t2d/cg2plain.py:"
                                      "11-49, FP64)" % it_cpu}})


def bench_kubc3d(a):
    """48^3 KUBC, nine right-hand sides in one batched FP32 PCG + FP64 refinement to 1e-8."""
    import torch
    import lpm_ops as ops
    import _lpmlib as L
    import synth
    import workloads
    dev = torch.device("cuda:0")
    P = workloads.kubc3d(48, a.seed, dev)
    dims = P["dims"]
    N = 48 ** 3

    def step(graph=True):
        b = -ops.apply(dims, 1.0, P["up"], P["kp64"], P["tp64"], layout=L.PACKED) * P["free"][:, None]
        M = ops.block_jacobi(dims, 1.0, P["kp64"], P["tp64"], free=P["free"], out_dtype=torch.float32,
                             layout=L.PACKED)
        return ops.pcg_refined(dims, 1.0, b, P["kp64"], P["tp64"], P["kp32"], P["tp32"], minv32=M, free=P["free"],
                               tol=a.tol, maxit=100000, inner_tol=1e-4, check=25, graph=graph, layout64=L.PACKED,
                               layout32=L.PACKED)
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        x, info = step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    # K u launch times from one more solve without graph replays (HIP events around each launch)
    L.lib().lpm_profile_ku(1)
    step(graph=False)
    torch.cuda.synchronize()
    ms = ctypes.c_double()
    cnt = ctypes.c_int64()
    L.lib().lpm_profile_ku_read(ctypes.byref(ms), ctypes.byref(cnt))
    L.lib().lpm_profile_ku(0)
    ku_ms = ms.value / max(cnt.value, 1)
    ku_bytes = N * (2 * 3 * 9 * 4 + 18 * 4)
    u = P["up"] + x
    sup = 0.0
    for j, (p, q) in enumerate(workloads.COMBINED):
        c = u[6 + j]
        sup = max(sup, float((c - u[p] - u[q]).abs().max() / c.abs().max()))
    # CPU: one right-hand side of the C block-Jacobi PCG for a bounded number of iterations
    lib = _oracle_lib()
    Kn, Tv = synth.directed_from_packed(dims, P["kp64"], P["tp64"])
    Mf = ops.full_from_sym(ops.block_jacobi(dims, 1.0, P["kp64"], P["tp64"], free=P["free"], layout=L.PACKED))
    b0 = (-ops.apply(dims, 1.0, P["up"][0], P["kp64"], P["tp64"], layout=L.PACKED) * P["free"][:, None])
    nbr = np.ascontiguousarray(P["lat"].nbr.astype(np.int64))
    arrs = [np.ascontiguousarray(t.cpu().numpy()) for t in (Kn, Tv, Mf, b0)]
    fr = np.ascontiguousarray(P["fixed"] == False).astype(np.uint8)  # noqa: E712
    xx = np.zeros_like(arrs[3])
    rel = np.zeros(1)
    pt = lambda v: v.ctypes.data_as(ctypes.c_void_p)
    its = 200
    t0 = time.perf_counter()
    done = lib.orc_pcg3(N, pt(nbr), pt(arrs[0]), pt(arrs[1]), pt(arrs[2]), pt(fr), pt(arrs[3]), pt(xx), 0.0, its,
                        pt(rel))
    t_it = (time.perf_counter() - t0) / max(done, 1)
    per_rhs_iters = info.iters.max()
    t_cpu = t_it * float(info.iters.sum())
    lat_ms = np.array(lat) * 1e3
    _line(METRIC, 9 * a.steps / sum(lat), "RHS solves/s", a.steps, a.warmup, float(lat_ms.mean()),
          "f32 inner + f64 refinement",
          "SC LPM 48^3, spheres phi 0.25 contrast 10, 9 KUBC right-hand sides in one batched block-Jacobi PCG, "
          "FP32 + FP64 refinement to rel-res %g" % a.tol,
          {"latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
           "solve": {"inner_iters_max": int(per_rhs_iters), "outer": int(np.max(info.outer)),
                     "relres_fp64": float(info.relmax), "superposition_rel_err": sup,
                     "phase_fraction": P["phase_fraction"]},
           "ku": {"avg_us": ku_ms * 1e3, "launches": cnt.value, "alg_bytes_per_launch": ku_bytes,
                  "achieved_GBs": ku_bytes / (ku_ms * 1e-3) / 1e9,
                  "note": "31.9 MB per batched K u: cache resident, reported as time"},
           "cpu_baseline": {"value": 9.0 / t_cpu, "unit": "RHS solves/s", "cores": int(lib.orc_threads()),
                            "kind": "port", "sample": "%d block-Jacobi PCG iterations of one RHS (FP64 C/OpenMP "
                            "restatement of coarse3d/lin3d.py:464-518), %.4f s/iter, scaled by the GPU's total "
                            "inner iterations over the 9 RHS" % (done, t_it)}})


def bench_amg256(a):
    """The 256^3 cement-like problem of the default line solved by AMG-PCG (timing/t3d/amg3d.py with AMG_CFG of
    timing/t3d/problems.py:19: 3x3x3 rigid-body-mode aggregation, W(1,1) cycles, dense coarsest level) in FP64 to
    rel-res tol.  The timed solve includes the numeric setup (block Jacobi, rho by power iteration, Galerkin
    stencils by probing, coarsest inverse), as the reference's timed AMG solves do; the aggregation geometry is
    built once per lattice and BC set, as in the reference."""
    import types
    import torch
    import amg3d
    dev = torch.device("cuda:0")
    P = build_problem(a.n, a.seed, dev)
    ops, L = P["ops"], P["L"]
    dims = P["dims"]
    n = a.n
    N = n ** 3
    b = -ops.apply(dims, 1.0, P["up"], P["kp64"], P["tp64"], layout=L.PACKED) * P["free"][:, None]
    lat = types.SimpleNamespace(nx=n, ny=n, nz=n, h=1.0, full=True, device=dev)
    t0 = time.perf_counter()
    cfg = dict(amg3d.AMG_CFG)
    geom = amg3d.AMGGeom(lat, (P["free"] > 0).cpu().numpy(), agg=cfg["agg"], coarse_max=cfg["coarse_max"])
    t_geom = time.perf_counter() - t0
    print("[bench] amg geometry %.1f s: %s" % (t_geom, geom.describe()), file=sys.stderr, flush=True)

    def step(fp64=False):
        ts = time.perf_counter()
        amg = amg3d.AMG3D(geom, P["kp64"] if fp64 else P["kp32"], P["tp64"] if fp64 else P["tp32"], nu=cfg["nu"],
                          gamma=cfg["gamma"], power_its=cfg["power_its"], starts="device",
                          dtype=torch.float64 if fp64 else torch.float32)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - ts
        if fp64:
            x, it, rr = amg.pcg(b, a.tol, 5000)
        else:
            x, it, _, rr = amg.pcg_refined(b, P["kp64"], P["tp64"], tol=a.tol, inner_tol=a.inner_tol, maxit=1000)
        torch.cuda.synchronize()
        return x, it, rr, t_setup, time.perf_counter() - ts, amg

    for w in range(a.warmup):
        _, it, rr, ts_, tt, _ = step()
        print("[bench] warmup %d: %d its, relres %.2e, setup %.3f s, total %.3f s" % (w, it, rr, ts_, tt),
              file=sys.stderr, flush=True)
    lat_s, its, setups = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(a.steps):
        x, it, rr, ts_, tt, amg = step()
        lat_s.append(tt)
        its.append(it)
        setups.append(ts_)
        print("[bench] step %d: %.3f s (setup %.3f s), %d its, relres %.2e" % (s, tt, ts_, it, rr), file=sys.stderr,
              flush=True)
    total = time.perf_counter() - t0
    # the same solve with the FP64 hierarchy and FP64 PCG (the reference's arithmetic), its true residual, and the
    # distance of both solutions to the block-Jacobi refined solve
    x64, it64, _, setup64, t64, amg64 = step(fp64=True)
    r = b + amg64._kff(x64)                    # _kff(x) = -K_ff x
    rel_true = float(torch.linalg.vector_norm(r * P["free"][:, None]) / torch.linalg.vector_norm(b))
    xb, info = solve(P, a.tol)
    dx64 = float(torch.linalg.vector_norm(x64 - xb) / torch.linalg.vector_norm(xb))
    dx = float(torch.linalg.vector_norm(x - xb) / torch.linalg.vector_norm(xb))
    lat_ms = np.array(lat_s) * 1e3
    out = {
        "metric": METRIC, "value": a.steps / total, "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": total / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 inner + f64 refinement", "data": "synthetic",
        "config": {"workload": "SC LPM %d^3 (%d DOF), 4-phase cement-like, tension z, AMG-PCG (3x3x3 rigid-body-mode "
                               "aggregation, W(1,1)) FP32 inside FP64 refinement to rel-res %g, numeric setup inside "
                               "the timed solve"
                               % (n, 3 * N, a.tol), "n": n, "tol": a.tol, "seed": a.seed, "amg": cfg, "inner_tol": a.inner_tol,
                   "levels": geom.describe()},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                       "max": float(lat_ms.max())},
        "solve": {"pcg_iters": int(np.median(its)), "setup_s": float(np.median(setups)),
                  "ms_per_iter": float((np.median(lat_s) - np.median(setups)) / max(1, np.median(its)) * 1e3),
                  "relres_true_fp64": rr, "rel_diff_vs_block_jacobi_solve": dx, "block_jacobi_iters": info.it,
                  "geometry_s": t_geom},
        "fp64_amg_pcg": {"s_per_solve": t64, "setup_s": setup64, "pcg_iters": it64, "relres_true_fp64": rel_true,
                         "rel_diff_vs_block_jacobi_solve": dx64},
    }
    if not a.no_cpu:
        print("[bench] cpu AMG baseline ...", file=sys.stderr, flush=True)
        del x, x64, xb, r, amg64
        torch.cuda.empty_cache()
        out["cpu_baseline"] = amg_cpu_baseline(P, it64)
    print(json.dumps(out), flush=True)


def bench_patch(a):
    """Local patch homogenisation (coarse3d/patch3d.py) of a 96^3 two-phase voxel field (contrast 5, fraction 0.35,
    seeded), s = 2: every coarse cell's (3s)^3 window solved for the 6 unit strains by the batched FP64 PCG and its
    C_ab from the batched particle energies.  value = homogenised cells per second; s = 3 reported beside it."""
    import torch
    import patch3d
    rng = np.random.default_rng(a.seed)
    n = 96
    E = np.where(rng.random(n ** 3) < 0.35, 5.0, 1.0)
    lam, mu = E * 0.2 / (1.2 * 0.6), 0.5 * E / 1.2
    Cv = np.stack([lam + 2 * mu, lam, mu], 1)
    res = {}
    for s, chunk in ((2, 8192), (3, 2048)):
        for _ in range(a.warmup):
            patch3d.patch_C(Cv, n, s, chunk=chunk)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            C = patch3d.patch_C(Cv, n, s, chunk=chunk)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        res[s] = dict(cells=len(C), s_per_field=dt, cells_per_s=len(C) / dt,
                      min_eig=float(np.linalg.eigvalsh(C).min()))
        print("[bench] patch s=%d: %d cells in %.3f s" % (s, len(C), dt), file=sys.stderr, flush=True)
    out = {"metric": "patch homogenisation cells/s", "value": res[2]["cells_per_s"], "unit": "cells/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": res[2]["s_per_field"] * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "patch homogenisation of a 96^3 two-phase voxel field (contrast 5, phi 0.35), "
                                  "s = 2 windows of 6^3 particles, 6 unit strains, batched FP64 PCG to 1e-13",
                      "n": n, "seed": a.seed}, "s2": res[2], "s3": res[3]}
    print(json.dumps(out), flush=True)


# HBM bytes per launch of the face-diagonal heads at n = 256 (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, KiB per
# dispatch, scaled by what the same counters read for a copy kernel of known size: x2.0 and x1.0)
HEADS_DIAG_TRAFFIC_256 = 4591490 * 1024 * 2.0 + 1399950 * 1024 * 1.0


def _bondnet_setup(n, s, seed):
    """Two-phase voxel field (fraction 0.4, contrast 4, nu 0.3 / 0.2, seeded), its voxel features and per-cell Voigt
    parameters, and a bond network of the reference architecture (ch 24, D 12, hid 64) with random weights (last
    layer N(0, 0.05), so that the outputs are not the Voigt values)."""
    import torch
    import geom3d as G
    import net3d
    rng = np.random.default_rng(seed)
    phase = (rng.random(n ** 3) < 0.4).astype(np.int64)
    X = net3d.voxel_features(phase, 4.0, 0.3, n, s)
    Cv = G.iso_C(np.where(phase == 1, 4.0, 1.0), np.where(phase == 1, 0.3, 0.2))
    torch.manual_seed(seed)
    net = net3d.BondNet(s)
    with torch.no_grad():
        for _, _, l3 in net.heads:
            l3.weight.normal_(0.0, 0.05)
            l3.bias.normal_(0.0, 0.05)
    return net, X, Cv, phase


def bondnet_cpu_baseline(s, seed, n_sample=48):
    """The oracle forward + post (NumPy with BLAS on the host's cores) on an n_sample^3 field of the same kind;
    the caller scales its time by the bond count."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
    import bondnet_ref as BR
    net, X, Cv, _ = _bondnet_setup(n_sample, s, seed)
    P = {k: v.detach().double().numpy() for k, v in net.state_dict().items() if not k.endswith("masks")}
    nc = n_sample // s
    cp = BR.voigt_cell_params(Cv, n_sample, s)
    t0 = time.perf_counter()
    outs, groups = BR.forward(X[None], nc, s, P)
    BR.post([o[0] for o in outs], groups, cp, nc)
    dt = time.perf_counter() - t0
    return dt, sum(len(g["bonds"]) for g in groups)


def _infer_split(net, Xd, cp, nc, s, Kn32, steps):
    """The same inference with the opt-in split-BF16 first layer (BondNet(precision="bf16x3"), same weights):
    time per inference, per-group head launches and the largest deviation of its Kn from the FP32 path's."""
    import torch
    import net3d
    sp = net3d.BondNet(s, precision="bf16x3")
    sp.load_state_dict(net.state_dict())
    sp = sp.to(Xd.device)
    Kn, _ = sp.bond_params(Xd, nc, cp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        Kn, _ = sp.bond_params(Xd, nc, cp)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    Fv, Fc = sp.features(Xd)
    evs = []
    sp.heads_run(Fv, Fc, nc, events=evs)
    torch.cuda.synchronize()
    dev_rel = float((Kn - Kn32).abs().max() / Kn32.abs().max())
    print("[bench] split-BF16 heads: %.1f ms per inference, Kn within %.1e of FP32" % (dt * 1e3, dev_rel),
          file=sys.stderr, flush=True)
    return {"value": 1.0 / dt, "unit": "inferences/s", "ms_per_step": dt * 1e3,
            "heads_axial_ms": evs[0][0].elapsed_time(evs[0][1]), "heads_diagonal_ms": evs[1][0].elapsed_time(evs[1][1]),
            "max_rel_dev_kn_vs_fp32": dev_rel,
            "note": "opt-in: first head layer as three BF16 matrix-core products of the hi/lo split operands, FP32 "
                    "accumulation; not the headline value"}


def _infer_then_solve(net, Xd, cp, nc, s, tol, reps=2):
    """Inference fused in front of the solve: Kn / Tv of the nc^3 coarse lattice (spacing h = s) from the network,
    packed, then tension z (planes z = 0 and z = nc - 1 fixed, top displaced by 1e-3 of the height) solved by
    block-Jacobi PCG in FP32 inside FP64 refinement to rel-res tol.  Not part of `value`: reported beside it."""
    import torch
    import _lpmlib as L
    import lpm_ops as ops
    dev = Xd.device
    dims, h = (nc, nc, nc), float(s)
    z = torch.arange(nc, device=dev).repeat_interleave(nc * nc)
    free = (~((z == 0) | (z == nc - 1))).to(torch.float64)
    up = torch.zeros((nc ** 3, 3), device=dev, dtype=torch.float64)
    up[z == nc - 1, 2] = 1e-3 * (nc - 1) * h
    res = []
    for r in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Kn, Tv = net.bond_params(Xd, nc, cp)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        kp64, tp64, nasym = ops.pack(dims, Kn[0], Tv[0])
        kp32, tp32 = kp64.float(), tp64.float()
        b = -ops.apply(dims, h, up, kp64, tp64, layout=L.PACKED) * free[:, None]
        M = ops.block_jacobi(dims, h, kp64, tp64, free=free, out_dtype=torch.float32, layout=L.PACKED)
        x, info = ops.pcg_refined(dims, h, b, kp64, tp64, kp32, tp32, minv32=M, free=free, tol=tol, maxit=100000,
                                  inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        res.append(dict(infer_ms=(t1 - t0) * 1e3, solve_ms=(t2 - t1) * 1e3, total_ms=(t2 - t0) * 1e3,
                        pcg_iters=int(info.it), outer=int(info.outer), relres_fp64=float(info.relmax),
                        asymmetric_bonds=int(nasym.sum())))
        print("[bench] infer + solve %d: %.1f + %.1f ms, %d its, relres %.2e" % (
            r, res[-1]["infer_ms"], res[-1]["solve_ms"], info.it, info.relmax), file=sys.stderr, flush=True)
    best = min(res, key=lambda d: d["total_ms"])
    best["workload"] = ("inferred Kn/Tv of the %d^3 coarse lattice -> tension z, block-Jacobi PCG FP32 + FP64 "
                        "refinement to rel-res %g" % (nc, tol))
    return best


def bench_infer(a):
    """Bond-network inference (coarse3d/net3d.py + the post of timing/t3d/infer_fast.py): a 256^3 two-phase voxel
    field, s = 2 -> the Kn / Tv (128^3, 18) of the coarse lattice.  The voxel features and per-cell Voigt parameters
    are on the GPU before timing; a step is trunk (3 IsoConv3d + 1x1), cell means, both head groups and the
    scatter.  Roofline of the dominant launch (the face-diagonal heads): first-layer FLOPs 2 nb H nin hid on the
    FP32 matrix cores."""
    import torch
    import net3d
    n, s = a.n, 2
    nc = n // s
    dev = torch.device("cuda:0")
    net, X, Cv, _ = _bondnet_setup(n, s, a.seed)
    net = net.to(dev)
    Xd = torch.as_tensor(X[None], device=dev)
    cp = net3d.voigt_cell_params(torch.as_tensor(Cv, device=dev), n, s)
    for w in range(a.warmup):
        Kn, Tv = net.bond_params(Xd, nc, cp)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        Kn, Tv = net.bond_params(Xd, nc, cp)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        print("[bench] infer step: %.1f ms" % (lat[-1] * 1e3), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    # per-launch timing (HIP events on the launching stream)
    ev_feat, ev_heads = [], []
    for _ in range(2):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        Fv, Fc = net.features(Xd)
        e1.record()
        ev_feat.append((e0, e1))
        Kn = torch.zeros((1, nc ** 3, 18), device=dev, dtype=torch.float64)
        Tv = torch.zeros_like(Kn)
        evs = []
        net.heads_run(Fv, Fc, nc, cp=cp[None].contiguous(), Kn=Kn, Tv=Tv, events=evs)
        ev_heads.append(evs)
        del Fv, Fc
    torch.cuda.synchronize()
    feat_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev_feat]))
    head_ms = [float(np.mean([ev[gi][0].elapsed_time(ev[gi][1]) for ev in ev_heads])) for gi in range(2)]
    D, hid = net.D, net.heads[0][0].weight.shape[0]
    groups = net.tables.groups
    nbs = [net._device_tables(nc, dev)[gi]["nb"] for gi in range(2)]
    flops = [2.0 * nbs[gi] * len(groups[gi]["perms"]) * (groups[gi]["nv"] + groups[gi]["ncell"]) * D * hid
             for gi in range(2)]
    peak = 157.3
    ach = flops[1] / (head_ms[1] * 1e-3) / 1e12
    lat_ms = np.array(lat) * 1e3
    out = {"metric": "bond-network inferences/s (fine 256^3 voxels -> coarse 128^3 Kn/Tv)",
           "value": a.steps / total, "unit": "inferences/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": total / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "BondNet s=2 (ch 24, D 12, hid 64, random weights) on a %d^3 two-phase voxel field "
                                  "(phi 0.4, contrast 4) -> Kn/Tv of the %d^3 coarse lattice" % (n, nc),
                      "n": n, "s": s, "seed": a.seed, "bonds": nbs},
           "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
           "stages_ms": {"trunk_and_pool": feat_ms, "heads_axial": head_ms[0], "heads_diagonal": head_ms[1]},
           "roofline": {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                        "traffic": HEADS_DIAG_TRAFFIC_256 if n == 256 else None,
                        "kernel": "bond_heads_kernel<8,4,3> (face-diagonal group)",
                        "alg_flops_per_launch": flops[1], "avg_us": head_ms[1] * 1e3,
                        "axial": {"alg_flops_per_launch": flops[0], "avg_us": head_ms[0] * 1e3,
                                  "achieved": flops[0] / (head_ms[0] * 1e-3) / 1e12}}}
    out["split_bf16"] = _infer_split(net, Xd, cp, nc, s, Kn, a.steps)
    out["pipeline"] = _infer_then_solve(net, Xd, cp, nc, s, a.tol, reps=2)
    if not a.no_cpu:
        print("[bench] cpu baseline (oracle forward on a 48^3 field) ...", file=sys.stderr, flush=True)
        dt, nb_s = bondnet_cpu_baseline(s, a.seed)
        scale = sum(nbs) / nb_s
        out["cpu_baseline"] = {"value": 1.0 / (dt * scale), "unit": "inferences/s",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
                               "sample": "oracle forward + post (NumPy, BLAS) of a 48^3 field (nc 24, %d bonds) in "
                                         "%.2f s, scaled by the bond count (x%.1f)" % (nb_s, dt, scale)}
    print(json.dumps(out), flush=True)


def bench_porous256(a):
    """256^3 cement-like microstructure with the pore particles removed (bond words), tension z, FP32 PCG + FP64
    refinement to 1e-8; the masked K p kernel's roofline and the C oracle's PCG on the same reduced lattice."""
    import torch
    import lpm_ops as ops
    import _lpmlib as L
    import workloads
    dev = torch.device("cuda:0")
    n = a.n
    t0 = time.perf_counter()
    P = workloads.porous3d(n, a.seed, dev)
    dims, w = P["dims"], P["words"]
    N = n ** 3
    print("[bench] porous %d^3 ready: solid %.3f, %.1f s" % (n, P["solid_fraction"], time.perf_counter() - t0),
          file=sys.stderr, flush=True)

    def step():
        b = -ops.apply(dims, 1.0, P["up"], P["kp64"], P["tp64"], layout=L.PACKED, bonds=w) * P["free"][:, None]
        M = ops.block_jacobi(dims, 1.0, P["kp64"], P["tp64"], free=P["free"], out_dtype=torch.float32,
                             layout=L.PACKED, bonds=w)
        return ops.pcg_refined(dims, 1.0, b, P["kp64"], P["tp64"], P["kp32"], P["tp32"], minv32=M, free=P["free"],
                               tol=a.tol, maxit=100000, inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED,
                               bonds=w)
    for wi in range(a.warmup):
        _, info = step()
        torch.cuda.synchronize()
        print("[bench] warmup %d: %d its, %d outer, relres %.2e" % (wi, info.it, info.outer, info.relmax),
              file=sys.stderr, flush=True)
    L.lib().lpm_profile_ku(1)
    lat, its = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s_ in range(a.steps):
        ts = time.perf_counter()
        x, info = step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        its.append(info.it)
        print("[bench] step %d: %.3f s, %d its, relres %.2e" % (s_, lat[-1], info.it, info.relmax), file=sys.stderr,
              flush=True)
    total = time.perf_counter() - t0
    ms = ctypes.c_double()
    cnt = ctypes.c_int64()
    L.lib().lpm_profile_ku_read(ctypes.byref(ms), ctypes.byref(cnt))
    L.lib().lpm_profile_ku(0)
    ku_ms = ms.value / max(cnt.value, 1)
    ku_bytes = N * (2 * 3 * 4 + 18 * 4 + 4)                 # u in + Ku out + 18 coefficients + bond word, FP32
    achieved = ku_bytes / (ku_ms * 1e-3) / 1e9
    lat_ms = np.array(lat) * 1e3
    extra = {"latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
             "solve": {"pcg_iters": int(np.median(its)), "outer": info.outer, "relres_fp64": info.relmax,
                       "solid_fraction": P["solid_fraction"], "removed_by_cleaning": P["removed_by_cleaning"]},
             "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                          "kernel": "ku3d_kernel<float,float,PACKED,64,4,MASKDOT,BONDS>", "launches": cnt.value,
                          "avg_us": ku_ms * 1e3, "alg_bytes_per_launch": ku_bytes}}
    if not a.no_cpu:
        extra["cpu_baseline"] = _porous_cpu(P, int(np.median(its)), a.cpu_budget)
        extra["fp64_oracle_64cube"] = _porous_subvolume(P["lab"], n, dev)
    _line(METRIC, a.steps / total, "solves/s", a.steps, a.warmup, total / a.steps * 1e3,
          "f32 inner + f64 refinement",
          "SC LPM %d^3 box, cement-like with pore particles removed (solid %.3f), tension z, block-Jacobi PCG FP32 "
          "+ FP64 refinement to rel-res %g" % (n, P["solid_fraction"], a.tol), extra)


def _masked_nbr(n, solid):
    """(N, 18) int64 neighbour table of the box with -1 where a slot leaves the box or either end is removed."""
    N = n ** 3
    p = np.arange(N, dtype=np.int64)
    ijk = np.stack([p % n, (p // n) % n, p // (n * n)], 1)
    nbr = np.empty((N, 18), np.int64)
    offs = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
                     (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1),
                     (0, 1, -1), (0, -1, 1)])
    for k, o in enumerate(offs):
        q = ijk + o
        ok = np.all((q >= 0) & (q < n), 1)
        qq = np.where(ok, (q[:, 2] * n + q[:, 1]) * n + q[:, 0], 0)
        nbr[:, k] = np.where(ok & solid & solid[qq], qq, -1)
    return nbr


def _porous_cpu(P, gpu_iters, budget_s):
    """C oracle block-Jacobi PCG (FP64, OpenMP) on the same reduced lattice for a bounded number of iterations."""
    import lpm_ops as ops
    import _lpmlib as L
    import synth
    lib = _oracle_lib()
    dims, w = P["dims"], P["words"]
    n = dims[0]
    N = n ** 3
    Kn, Tv = synth.directed_from_packed(dims, P["kp64"], P["tp64"])
    Kn = np.ascontiguousarray(Kn.cpu().numpy())
    Tv = np.ascontiguousarray(Tv.cpu().numpy())
    M = ops.full_from_sym(ops.block_jacobi(dims, 1.0, P["kp64"], P["tp64"], free=P["free"], layout=L.PACKED,
                                           bonds=w))
    M = np.ascontiguousarray(M.cpu().numpy())
    b = -ops.apply(dims, 1.0, P["up"], P["kp64"], P["tp64"], layout=L.PACKED, bonds=w) * P["free"][:, None]
    b = np.ascontiguousarray(b.cpu().numpy())
    fr = np.ascontiguousarray((P["free"] > 0).cpu().numpy().astype(np.uint8))
    nbr = _masked_nbr(n, P["solid"])
    ptr = lambda v: v.ctypes.data_as(ctypes.c_void_p)
    rel = np.zeros(1)
    x = np.zeros_like(b)
    t0 = time.perf_counter()
    lib.orc_pcg3(N, ptr(nbr), ptr(Kn), ptr(Tv), ptr(M), ptr(fr), ptr(b), ptr(x), 0.0, 2, ptr(rel))
    t_probe = (time.perf_counter() - t0) / 2
    its = int(max(3, min(200, budget_s / max(t_probe, 1e-6))))
    x[:] = 0
    t0 = time.perf_counter()
    done = lib.orc_pcg3(N, ptr(nbr), ptr(Kn), ptr(Tv), ptr(M), ptr(fr), ptr(b), ptr(x), 0.0, its, ptr(rel))
    t_it = (time.perf_counter() - t0) / max(done, 1)
    return {"value": 1.0 / (t_it * gpu_iters), "unit": "solves/s", "cores": int(lib.orc_threads()), "kind": "port",
            "sample": "%d block-Jacobi PCG iterations (FP64, C/OpenMP restatement of coarse3d/lin3d.py:464-518) on "
                      "the same reduced %d^3 lattice: %.3f s per iteration, scaled by the GPU solve's %d PCG "
                      "iterations" % (done, n, t_it, gpu_iters)}


def _porous_subvolume(lab, n_full, dev, n=64):
    """GPU (bond words, FP32 + FP64 refinement to 1e-8) vs the C oracle's FP64 PCG to 1e-13 on the reduced lattice
    (compact numbering) of the n^3 corner of the same labels, its solid set cleaned again."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lpm_ref as R
    import lpm_ops as ops
    import _lpmlib as L
    import workloads
    sub = lab.reshape(n_full, n_full, n_full)[:n, :n, :n].ravel()
    P = workloads.porous3d(n, 0, dev, lab=sub)
    dims, w = P["dims"], P["words"]
    b = -ops.apply(dims, 1.0, P["up"], P["kp64"], P["tp64"], layout=L.PACKED, bonds=w) * P["free"][:, None]
    M = ops.block_jacobi(dims, 1.0, P["kp64"], P["tp64"], free=P["free"], out_dtype=torch.float32, layout=L.PACKED,
                         bonds=w)
    x, info = ops.pcg_refined(dims, 1.0, b, P["kp64"], P["tp64"], P["kp32"], P["tp32"], minv32=M, free=P["free"],
                              tol=1e-8, maxit=100000, inner_tol=1e-4, layout64=L.PACKED, layout32=L.PACKED, bonds=w)
    ug = (P["up"] + x).cpu().numpy()
    lat = R.ActiveBox(dims, P["solid"])
    Kf, Tf = R.two_end_mean(R.Box(dims), P["ppar"])
    Kn, Tv = Kf[lat.sites], Tf[lat.sites]
    fixed = P["fixed"].cpu().numpy()[lat.sites]
    up = P["up"].cpu().numpy()[lat.sites]
    bo = np.ascontiguousarray(-R.apply(lat, up, Kn, Tv) * (~fixed)[:, None])
    Mo = np.ascontiguousarray(R.precond(lat, Kn, Tv, ~fixed))
    lib = _oracle_lib()
    pt = lambda v: v.ctypes.data_as(ctypes.c_void_p)
    xo = np.zeros_like(bo)
    rel = np.zeros(1)
    it = lib.orc_pcg3(lat.N, pt(np.ascontiguousarray(lat.nbr.astype(np.int64))), pt(np.ascontiguousarray(Kn)),
                      pt(np.ascontiguousarray(Tv)), pt(Mo), pt(np.ascontiguousarray((~fixed).astype(np.uint8))),
                      pt(bo), pt(xo), 1e-13, 100000, pt(rel))
    uo = up + xo
    ugc = ug[lat.sites]
    so, sg = R.stress(lat, uo, Kn, Tv), R.stress(lat, ugc, Kn, Tv)
    top = lat.ijk[:, 2] == n - 1
    Fo, Fg = R.apply(lat, uo, Kn, Tv)[top, 2].sum(), R.apply(lat, ugc, Kn, Tv)[top, 2].sum()
    return {"subvolume": "%d^3 corner, pores removed, tension z" % n, "solid_particles": int(lat.N),
            "gpu_relres_fp64": float(info.relmax), "oracle_iters": int(it), "oracle_relres": float(rel[0]),
            "rel_err_u": float(np.abs(ugc - uo).max() / np.abs(uo).max()),
            "rel_err_stress": float(np.abs(sg - so).max() / np.abs(so).max()),
            "rel_err_reaction": float(abs(Fg - Fo) / abs(Fo))}


def bench_serve(a):
    """Continuous batching of 512 coarse 2D requests (24^2-144^2, FP32 PCG + FP64 refinement to 1e-8)."""
    import torch
    import serve
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    reqs = serve.make_requests(a.requests, a.seed, dev)
    print("[serve] %d requests built in %.1f s" % (len(reqs), time.perf_counter() - t0), file=sys.stderr, flush=True)
    # plain per-solve GPU CG on a sample
    nplain = min(a.plain_sample, len(reqs))
    t0 = time.perf_counter()
    for r in reqs[:nplain]:
        serve.plain_cg(r, a.tol)
    torch.cuda.synchronize()
    plain_rate = nplain / (time.perf_counter() - t0)
    print("[serve] plain GPU CG: %.1f solves/s on %d requests" % (plain_rate, nplain), file=sys.stderr, flush=True)
    srv = serve.Server(max_slots=a.slots, device=dev, budget=a.budget, threads=a.threads)
    serve.run_stream(srv, reqs[:min(64, len(reqs))])                       # warm-up (graph capture, clocks)
    sat = serve.run_stream(srv, reqs)
    print("[serve] saturation: %.1f solves/s" % sat["solves_per_s"], file=sys.stderr, flush=True)
    rate = a.load * sat["solves_per_s"]
    arr = serve.poisson_arrivals(len(reqs), rate, a.seed + 1)
    pois = serve.run_stream(srv, reqs, arr)
    print("[serve] poisson %.1f/s: p50 %.2f ms p99 %.2f ms" % (rate, pois["p50_ms"], pois["p99_ms"]),
          file=sys.stderr, flush=True)
    srv.close()
    sizes = [r.nx for r in reqs]
    _line(METRIC, sat["solves_per_s"], "solves/s", len(reqs), 64, sat["wall_s"] / len(reqs) * 1e3,
          "f32 inner + f64 refinement",
          "serving: %d coarse 2D requests %d^2-%d^2, FP32 PCG + FP64 refinement to rel-res %g, continuous batching"
          % (len(reqs), min(sizes), max(sizes), a.tol),
          {"latency_ms": {"p50": pois["p50_ms"], "p99": pois["p99_ms"], "max": pois["max_ms"],
                          "arrival_rate_per_s": rate, "load": a.load},
           "saturation": sat, "poisson": pois,
           "plain_gpu_cg": {"solves_per_s": plain_rate, "sample": nplain,
                            "speedup": sat["solves_per_s"] / plain_rate},
           "slots": a.slots, "budget": a.budget, "threads": a.threads})


def bench_train3d(a):
    """Differentiable coarse 3D training step: 16 specimens x 9 loads on 24^3, forward PCG + adjoint + Adam."""
    import torch
    import torch_lpm3d as T3
    import workloads
    dev = torch.device("cuda:0")
    W = workloads.train3d(16, 48, 2, a.seed, dev)
    model = T3.CoarseT(W["nc"], W["s"], device=dev, dtype=torch.float64, mixed=True)
    Kn0, Tv0 = W["Kn"].double(), W["Tv"].double()
    sym = T3.SymScatter(model.lat, dev)
    theta = torch.zeros((16, model.N, 9), device=dev, dtype=torch.float64, requires_grad=True)
    with torch.no_grad():
        g = torch.Generator(device="cpu").manual_seed(a.seed)
        pert = sym(0.1 * torch.randn((16, model.N, 9), generator=g, dtype=torch.float64).to(dev))
        tgt = model.run(Kn0 * torch.exp(pert), Tv0, tol=a.tol, differentiable=False)
    opt = torch.optim.Adam([theta], lr=1e-2)

    def step():
        opt.zero_grad()
        Kn = Kn0 * torch.exp(sym(theta))
        o = model.run(Kn, Tv0, tol=a.tol)
        loss = ((o["u"] - tgt["u"]) ** 2).mean() / (tgt["u"] ** 2).mean() + \
            ((o["sig"] - tgt["sig"]) ** 2).mean() / (tgt["sig"] ** 2).mean()
        loss.backward()
        opt.step()
        return float(loss.detach())
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    lat, losses = [], []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        losses.append(step())
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    lat_ms = np.array(lat) * 1e3
    _line("training steps/s (16 specimens x 9 loads, forward + adjoint)", a.steps / sum(lat), "steps/s", a.steps,
          a.warmup, float(lat_ms.mean()), "f32 inner + f64 refinement",
          "coarse SC 24^3 x 16 specimens x 9 loads, forward refined PCG + adjoint PCG + gradient kernels + Adam",
          {"latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
           "loss": losses, "solves_per_step": 2 * 16 * 9})



def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--probe", type=int, default=0, help="run the profiler probe on an n^3 lattice and exit")
    ap.add_argument("--probe-reps", type=int, default=10)
    ap.add_argument("--config", default="cube256",
                    choices=["cube256", "plain2d", "kubc3d", "serve", "train3d", "porous256", "amg256", "patch",
                             "infer"])
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--slots", type=int, default=512)
    ap.add_argument("--budget", type=int, default=400000)
    ap.add_argument("--load", type=float, default=0.8)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--threads", type=int, default=1024, help="serving: workgroup size per request slot")
    ap.add_argument("--inner-tol", type=float, default=1e-6, help="amg256: relative tolerance of the FP32 inner solves")
    a = ap.parse_args()
    if a.probe:
        probe(a.probe, a.probe_reps)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solve path runs on a single MI355X (run with --gpus 1); sharding over "
              "several GPUs is a design, not an implementation, of this library", file=sys.stderr)
        sys.exit(2)
    if a.config != "cube256":
        {"plain2d": bench_plain2d, "kubc3d": bench_kubc3d, "serve": bench_serve, "train3d": bench_train3d,
         "porous256": bench_porous256, "amg256": bench_amg256, "patch": bench_patch,
         "infer": bench_infer}[a.config](a)
        return
    import torch
    dev = torch.device("cuda:0")
    P = build_problem(a.n, a.seed, dev)
    L = P["L"]
    N = a.n ** 3
    print("[bench] problem ready: %d^3, phases %s, setup %.1f s" % (a.n, np.round(P["fractions"], 3), P["t_setup"]),
          file=sys.stderr, flush=True)
    for w in range(a.warmup):
        _, info = solve(P, a.tol)
        torch.cuda.synchronize()
        print("[bench] warmup %d: %d its, %d outer, relres %.2e" % (w, info.it, info.outer, info.relmax),
              file=sys.stderr, flush=True)
    lat = []
    L.lib().lpm_profile_ku(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for s in range(a.steps):
        ts = time.perf_counter()
        _, info = solve(P, a.tol)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        its.append(info.it)
        print("[bench] step %d: %.3f s, %d its, %d outer, relres %.2e" % (s, lat[-1], info.it, info.outer,
                                                                         info.relmax), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    ms = ctypes.c_double()
    cnt = ctypes.c_int64()
    L.lib().lpm_profile_ku_read(ctypes.byref(ms), ctypes.byref(cnt))
    L.lib().lpm_profile_ku(0)
    ku_ms = ms.value / max(cnt.value, 1)
    ku_bytes = N * (2 * 3 * 4 + 18 * 4)
    achieved = ku_bytes / (ku_ms * 1e-3) / 1e9
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC, "value": a.steps / total, "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": total / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 inner + f64 refinement", "data": "synthetic",
        "config": {"workload": "SC LPM %d^3 (%d DOF), 4-phase cement-like, tension z, block-Jacobi PCG FP32 + FP64 "
                               "refinement to rel-res %g" % (a.n, 3 * N, a.tol), "n": a.n, "tol": a.tol,
                   "seed": a.seed},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                       "max": float(lat_ms.max())},
        "solve": {"pcg_iters": int(np.median(its)), "outer": info.outer, "relres_fp64": info.relmax},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": KU_TRAFFIC_256 if a.n == 256 else None,
                     "kernel": "ku3d_kernel<float,float,PACKED,64,4,MASKDOT>", "launches": cnt.value,
                     "avg_us": ku_ms * 1e3, "alg_bytes_per_launch": ku_bytes},
    }
    if not a.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(P, int(np.median(its)), a.cpu_budget, a.tol)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
