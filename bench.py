#!/usr/bin/env python3
"""SIMP iterations/s on the 120x40x40 hex8 cantilever, one MI355X.

A step is one SIMP iteration of the hot path on the device-resident problem:
density filter, SIMP densities + Jacobi diagonal, PCG solve to
||f - K u|| <= 1e-8 ||f||, compliance + sensitivity, filter pullback, OC
update.  The timed region is one SIMP run of --steps iterations (50, the
workload's OC iteration count) from the initial design rho = volfrac, with
the design and load already on the device.  Six paths are timed on the
same problem:
  fp64        FP64 Jacobi PCG (the reference's solver), warm-started from the
              previous displacement;
  fp64_mg     FP64 PCG preconditioned by the multigrid V-cycle (separates the
              preconditioner's gain from the FP32 arithmetic's);
  mixed       FP32 inner Jacobi PCG with FP64 residual replacement, previous-
              displacement warm start;
  mixed_mg    the mixed path with the FP32 PCG preconditioned by a geometric
              multigrid V-cycle instead of Jacobi;
  learned     the mixed path warm-started from the point of span{previous
              displacement, BF16 network's prediction} with the smallest residual;
  learned_mg  the learned warm start on the multigrid-preconditioned path.
`value` is the learned_mg path's iterations/s (FP32 + learned warm start, the
headline configuration; mixed_mg's if no trained weights are found).

Prints ONE JSON line (plus progress on stderr).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import _pkg  # noqa: E402

METRIC = "SIMP iter/s, 120×40×40 hex8: FP32+learned vs FP64 PCG at rel res 1e-8"
HBM_PEAK_GBS = 8000.0
WEIGHTS = [os.path.join(ROOT, "topopt.jl_amd", "weights", n) for n in
           ("warmstart_120x40x40.npz", "warmstart_64x32x32.npz")]


def default_weights(nels):
    """The network trained or fine-tuned on this grid, else the first of WEIGHTS."""
    own = os.path.join(ROOT, "topopt.jl_amd", "weights",
                       "warmstart_{}x{}x{}.npz".format(*nels))
    return next((w for w in [own] + WEIGHTS if os.path.exists(w)), None)


KERNEL_NAMES = {
    "pcg32": "k_matvec<float,3,PCG> (fused direction update + K p)",
    "update32": "k_pcg_update<float> (x += a p, r -= a q, partials)",
    "mg_pre": "k_mg_smooth<3,0> finest level (x = w D^-1 b, r = b - K x)",
    "mg_restrict": "k_mg_restrict<3> finest -> first coarse level",
    "mg_coarse": ("k_mg_dense_rows: coarsest level, x = M b with M the dense (symmetrised) "
                  "Chebyshev operator, one wave per row"),
    "mg_post": "k_mg_smooth<3,1> finest level (x += P xc, z = x + w D^-1 (b - K x))",
}


ROOFLINE_NOTES = {
    "mg_coarse": ("the coarsest multigrid level (a few thousand DOFs) applies its dense "
                  "Chebyshev operator: two small launches whose time is launch latency, not "
                  "traffic.  The grid-wide finest-level kernels' fractions are in kernel_gbs "
                  "(mg_post, mg_pre, pcg32)."),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_solver(T, nels, precision):
    prob = T.PointLoadCantilever(nels, load="centre")
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0),
                    precision=precision)
    return prob, s


def run_path(T, name, nels, precision, warm, steps, warmup, volfrac, rmin, rel_tol, inner_rel,
             weights=None, optimizer="oc"):
    import ctypes as C
    from topopt_jl_amd._lib import check, dptr
    prob, s = make_solver(T, nels, precision)
    if weights is not None:
        from topopt_jl_amd.nn import WarmStartNet
        s.attach_warm_start(WarmStartNet.for_problem(prob, weights))
    kw = dict(volfrac=volfrac, rmin=rmin, rel_tol=rel_tol, precision=precision,
              warm_start=warm, inner_rel_tol=inner_rel, optimizer=optimizer)
    t0 = time.perf_counter()
    w = T.simp_run(s, iterations=warmup, **kw)
    log(f"[{name}] warmup {warmup} it: {time.perf_counter() - t0:.2f}s pcg={list(w.pcg_iters)}")
    T.simp_run(s, iterations=0, **kw)  # initial design + load on the device, u = 0
    t0 = time.perf_counter()
    r = T.simp_run(s, iterations=steps, x0="device", **kw)
    dt = time.perf_counter() - t0
    log(f"[{name}] timed {steps} it: {dt:.3f}s pcg={list(r.pcg_iters)} obj={r.obj[-1]:.6g} "
        f"converged={r.converged}" + (f" network predictions {r.nn_calls}"
                                      if warm == "learned" else ""))
    xphys = np.empty(prob.ncells)
    check(s.lib.tohip_get_xphys(s.ctx.ptr, dptr(xphys)), s.ctx.ptr, "tohip_get_xphys")
    return dict(solver=s, prob=prob, seconds=dt, iters_per_s=steps / dt, obj=r.obj,
                pcg=r.pcg_iters, converged=r.converged, xphys=xphys, memory=s.memory(),
                nn_calls=r.nn_calls if warm == "learned" else None)


ALL_KERNELS = ("matvec64", "matvec32", "pcg64", "pcg32", "update64", "update32", "filter",
               "filter_t", "energy", "oc", "copy64", "mg_pre", "mg_restrict", "mg_coarse",
               "mg_post", "vcycle", "csr_spmv")


def kernel_times(T, s, prob, reps, rmin, names=ALL_KERNELS):
    """Mean kernel durations (us) from HIP events on the library stream, on the
    matvec micro-benchmark inputs: rho ~ U[0.01, 1], u ~ N(0, 1), seed 0."""
    import ctypes as C
    from topopt_jl_amd._lib import KERNELS, check
    g = np.random.default_rng(0)
    rho = g.uniform(0.01, 1.0, prob.ncells)
    u = g.standard_normal(prob.ndofs)
    u[prob.fixed.astype(bool)] = 0.0
    s.set_densities(rho)
    s.matvec(u)
    T.DensityFilterFun(s, rmin)
    out = {}
    for name in names:
        us = C.c_double()
        check(s.lib.tohip_bench_kernel(s.ctx.ptr, KERNELS[name], reps, C.byref(us)), s.ctx.ptr,
              "tohip_bench_kernel")
        out[name] = us.value
    return out


def mg_levels(nels, max_nodes=600, max_cells=400):
    """Cell counts of the multigrid levels: halve while every count is even
    until the coarsest fits the one-workgroup coarse kernel."""
    gs = [tuple(nels)]
    fits = lambda g: (np.prod([n + 1 for n in g]) <= max_nodes and np.prod(g) <= max_cells)
    while len(gs) < 2 or not fits(gs[-1]):
        if any(n % 2 for n in gs[-1]):
            break
        gs.append(tuple(n // 2 for n in gs[-1]))
    return gs


def algorithmic_bytes(prob):
    """Per-launch algorithmic HBM bytes of each kernel (s = scalar size)."""
    nd, nc = prob.ndofs, prob.ncells
    dim = len(prob.nels)
    lv = mg_levels(prob.nels)
    nn_of = lambda g: int(np.prod([n + 1 for n in g]))
    nd1, nn1 = dim * nn_of(lv[1]), nn_of(lv[1])
    ndL, nnL, ncL = dim * nn_of(lv[-1]), nn_of(lv[-1]), int(np.prod(lv[-1]))
    nn = nd // dim
    mv = lambda s: 2 * s * nd + s * nc  # read x, write y, read E once
    return {
        # finest-level multigrid kernels (FP32): read b, D^-1, E, mask; write x, r
        "mg_pre": 4 * 4 * nd + 4 * nc + nn,
        # read r (fine), write b and read the mask (first coarse level)
        "mg_restrict": 4 * nd + 4 * nd1 + nn1,
        # read b, D^-1, E, mask; write x (coarsest level, once per cycle)
        # dense coarsest solve: read M (ndof^2 floats) and b, write x
        "mg_coarse": 4 * ndL * ndL + 8 * ndL,
        # read x, b, D^-1, E, mask, the coarse correction; write z
        "mg_post": 4 * 4 * nd + 4 * nc + nn + 4 * nd1,
        "matvec64": mv(8), "matvec32": mv(4),
        "pcg64": mv(8) + 3 * 8 * nd, "pcg32": mv(4) + 3 * 4 * nd,  # + read r, dinv; write p
        "update64": 7 * 8 * nd, "update32": 7 * 4 * nd,  # x,r rw; p,q,dinv read
        "filter": 2 * 8 * nc, "filter_t": 2 * 8 * nc,
        "energy": 8 * nd + 3 * 8 * nc,  # read u, x~; write ce, grad
        "copy64": 16 * nd,  # plain copy: the counter-scaling reference
        # assembled K: every stored entry (value + column), the row pointers,
        # x read once, y written
        "csr_spmv": 12 * csr_nnz(prob.nels) + 8 * (nd + 1) + 16 * nd,
    }


def assembly_comparator(T, nels, rho, rel_tol):
    """One FP64 Jacobi-PCG solve of the same cantilever at uniform rho through
    the matrix-free operator and through the assembled CSR matrix
    (CGAssemblySolver), each after a warm-up solve; the assembled path
    re-assembles K inside every solve, as the reference's solver() does."""
    import ctypes as C
    from topopt_jl_amd._lib import check
    prob = T.PointLoadCantilever(nels, load="centre")
    nf = float(np.linalg.norm(prob.f))
    out = {}
    for name, tag, kw in (("matrix_free", T.HIPMatrixFreeSolver, {}),
                          ("assembled", T.HIPCGAssemblySolver, {"preconditioner": "jacobi"})):
        s = T.FEASolver(tag, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0),
                        cg_max_iter=200000, **kw)
        s.vars = np.full(prob.ncells, rho)
        s(abstol=rel_tol * nf, reltol=0.0)
        t0 = time.perf_counter()
        s(abstol=rel_tol * nf, reltol=0.0)
        dt = time.perf_counter() - t0
        r = dict(solve_s=round(dt, 4), pcg_iters=int(s.last_iters),
                 ms_per_iter=round(1e3 * dt / max(s.last_iters, 1), 4),
                 converged=bool(s.converged), u=s.u.copy())
        if name == "assembled":
            nnz, nbytes = C.c_int64(), C.c_int64()
            check(s.lib.tohip_csr_info(s.ctx.ptr, C.byref(nnz), C.byref(nbytes), None),
                  s.ctx.ptr, "tohip_csr_info")
            t0 = time.perf_counter()
            check(s.lib.tohip_csr_assemble(s.ctx.ptr, 0), s.ctx.ptr, "tohip_csr_assemble")
            r.update(nnz=nnz.value, matrix_bytes=nbytes.value,
                     assemble_s=round(time.perf_counter() - t0, 4))
        out[name] = r
        s.close()
    ua, um = out["assembled"].pop("u"), out["matrix_free"].pop("u")
    out["rel_diff_u"] = float(np.linalg.norm(ua - um) / np.linalg.norm(um))
    return out


def heat_stats(T, nels, reps, rel_tol):
    """Heat conduction on the same grid: flux 1 through the top (z = max)
    facets, T = 0 on the x = 0 nodes, conductivity rho(x) with x ~ U[0.01, 1]
    (seed 0), p = 3: the conduction operator's mean duration and one FP64
    Jacobi-PCG solve (after a warm-up solve)."""
    import ctypes as C
    prob = T.HeatConductionProblem(nels, Tleft=0.0, Tright=None, heatflux={"top": 1.0})
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-3, penalty=T.PowerPenaltyFun(3.0),
                    cg_max_iter=200000, abstol=0.0, reltol=rel_tol)
    s.set_densities(np.random.default_rng(0).uniform(0.01, 1.0, prob.ncells))
    s()
    t0 = time.perf_counter()
    s()
    dt = time.perf_counter() - t0
    us = C.c_double()
    rc = s.lib.tohip_heat_bench_matvec(s.ptr, reps, C.byref(us))
    if rc != 0:
        raise RuntimeError("tohip_heat_bench_matvec failed")
    nbytes = 16 * prob.nnodes + 8 * prob.ncells + prob.nnodes  # x, y; rho per cell; mask
    out = dict(nnodes=prob.nnodes, matvec_us=round(us.value, 3),
               matvec_gbs=round(nbytes / (us.value * 1e-6) / 1e9, 1),
               algorithmic_bytes=nbytes, solve_s=round(dt, 4), pcg_iters=int(s.last_iters),
               converged=bool(s.converged))
    s.close()
    return out


def csr_nnz(nels):
    """Stored entries of the assembled K: dim^2 per (node, neighbour) pair."""
    dim = len(nels)
    return dim * dim * int(np.prod([3 * (n + 1) - 2 for n in nels]))


def cpu_baseline(nels, volfrac, rmin, rel_tol, budget_s):
    """Oracle (C/OpenMP FP64 restatement of the reference algorithm: BFS-built
    sparse filter J, per-cell dense Ke products with the dof_cells gather,
    Jacobi PCG to ||r|| <= rel_tol ||f||) on the host cores.  Measured: the
    first two SIMP iterations of the run from rho = volfrac (filter, solve
    from zero then from the previous displacement, energies, pullback, OC)
    with every core.  The 1-thread figure is a bounded sample: PCG
    iterations and one filter/energy/pullback/OC pass at 1 thread, scaled by
    the PCG iterations those two iterations took."""
    O = _pkg.load_oracle()
    O.build()
    threads = O.set_threads(0)
    log(f"cpu baseline: building the oracle problem and filter ({threads} threads)")
    prob = O.Cantilever(tuple(nels), load="centre")
    orc = O.Oracle(prob, xmin=1e-9, p=3.0)
    flt = O.Filter(tuple(nels), rmin)
    nf = float(np.linalg.norm(prob.f))

    def simp_iteration(x, u0):
        t0 = time.perf_counter()
        xt = flt.apply(x)
        orc.set_vars(xt)
        u, it, _ = orc.pcg(prob.f, abstol=rel_tol * nf, reltol=0.0, maxiter=200000, x0=u0)
        obj, ce, g = orc.element_energy(u)
        dc = flt.apply_t(g)
        xn, _ = O.oc_update(x, dc, flt.colsum, volfrac, 0.2, 1e-12)
        t_pass = time.perf_counter() - t0
        return xn, u, it, t_pass

    x = np.full(prob.nel, volfrac)
    log("cpu baseline: timing SIMP iterations 1 and 2 (all cores)")
    x1, u1, it1, t1 = simp_iteration(x, None)
    x2, u2, it2, t2 = simp_iteration(x1, u1)
    log(f"cpu baseline: {t1:.2f} s ({it1} PCG iterations), {t2:.2f} s ({it2})")
    # 1 thread: PCG iterations and the rest of the pass, on the same problem
    O.set_threads(1)
    try:
        orc.set_vars(flt.apply(x1))
        t0 = time.perf_counter()
        orc.pcg(prob.f, abstol=0.0, reltol=0.0, maxiter=2)
        t_cal = (time.perf_counter() - t0) / 2
        n = int(max(5, min(2000, 0.6 * budget_s / max(t_cal, 1e-6))))
        log(f"cpu baseline: timing {n} Jacobi-PCG iterations at 1 thread")
        t0 = time.perf_counter()
        _, it, _ = orc.pcg(prob.f, abstol=0.0, reltol=0.0, maxiter=n)
        t_pcg1 = (time.perf_counter() - t0) / max(it, 1)
        t0 = time.perf_counter()
        xt = flt.apply(x1)
        orc.set_vars(xt)
        obj, ce, g = orc.element_energy(u1)
        dc = flt.apply_t(g)
        O.oc_update(x1, dc, flt.colsum, volfrac, 0.2, 1e-12)
        t_rest1 = time.perf_counter() - t0
    finally:
        O.set_threads(threads)
    per_iter1 = t_pcg1 * (it1 + it2) / 2 + t_rest1
    return {"value": 2.0 / (t1 + t2), "unit": "SIMP iter/s", "cores": threads, "kind": "port",
            "sample": (f"oracle C/OpenMP FP64 restatement of the reference algorithm, "
                       f"{nels[0]}x{nels[1]}x{nels[2]}: SIMP iterations 1 and 2 from "
                       f"rho={volfrac} measured end to end ({t1:.2f} s with {it1} Jacobi-PCG "
                       f"iterations from zero, {t2:.2f} s with {it2} from the previous "
                       f"displacement) on {threads} threads"),
            "single_thread": {"value": 1.0 / per_iter1, "cores": 1,
                              "sample": (f"{it} Jacobi-PCG iterations timed at 1 thread "
                                         f"({t_pcg1 * 1e3:.1f} ms/iter) + one filter/energy/"
                                         f"pullback/OC pass ({t_rest1 * 1e3:.0f} ms), scaled to "
                                         f"the {(it1 + it2) / 2:.0f} PCG iterations per SIMP "
                                         f"iteration measured above")}}


def summary(p, ref):
    out = {"iters_per_s": round(p["iters_per_s"], 4),
           "pcg_iters_total": int(np.sum(p["pcg"])),
           "pcg_iters": [int(v) for v in p["pcg"]],
           "converged": bool(p["converged"]),
           "solver_device_bytes": int(p["memory"]["solver_bytes"]),
           "device_used_bytes": int(p["memory"]["device_used_bytes"]),
           "final_obj": float(p["obj"][-1])}
    if p["nn_calls"] is not None:  # predictions used before the network stopped paying
        out["network_predictions"] = int(p["nn_calls"])
    if ref is not None:  # same optimizer as the FP64 path: compare trajectories
        out["max_rel_compliance_err_vs_fp64"] = float(np.max(np.abs(p["obj"] / ref["obj"] - 1)))
        out["final_rho_max_abs_diff_vs_fp64"] = float(np.max(np.abs(p["xphys"] - ref["xphys"])))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nels", type=str, default="120,40,40")
    ap.add_argument("--volfrac", type=float, default=0.3)
    ap.add_argument("--rmin", type=float, default=1.5)
    ap.add_argument("--rel-tol", type=float, default=1e-8)
    ap.add_argument("--reps", type=int, default=200)
    ap.add_argument("--inner-rel", type=float, default=None,
                    help="FP64 residual refresh ratio of the mixed paths (default 0.1 with "
                         "Jacobi, 1e-3 with multigrid)")
    ap.add_argument("--weights", default=None, help="warm-start network weights (.npz)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--nn-only", action="store_true",
                    help="only the warm-start network's inference on the workload grid (for "
                         "profiler runs): 3 + 20 predictions, prints the mean time as JSON")
    ap.add_argument("--kernels", default=None,
                    help="with --kernels-only: comma-separated subset of " + ",".join(ALL_KERNELS)
                         + " (no heat-conduction timing)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the kernel micro-benchmarks (for profiler runs); prints their "
                         "mean durations and algorithmic GB/s as one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got "
              f"--gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)
    nels = tuple(int(v) for v in args.nels.split(","))
    T = _pkg.load_package()
    from topopt_jl_amd._lib import device_info
    ndev, devname = device_info()
    if ndev < 1:
        print("no GPU visible to libtopopt_hip", file=sys.stderr)
        sys.exit(1)
    log(f"device: {devname}")
    if args.nn_only:
        from topopt_jl_amd.nn import WarmStartNet
        weights = args.weights or default_weights(nels)
        prob = T.PointLoadCantilever(nels, load="centre")
        net = WarmStartNet.for_problem(prob, weights)
        x = np.full(prob.ncells, args.volfrac)
        f = np.ascontiguousarray(prob.f)
        for _ in range(3):
            net.predict(x, f)
        t0 = time.perf_counter()
        for _ in range(20):
            net.predict(x, f)
        dt = (time.perf_counter() - t0) / 20
        print(json.dumps({"nels": list(nels), "weights": weights and os.path.relpath(weights, ROOT),
                          "predict_ms": round(dt * 1e3, 4), "note": "wall time per prediction, "
                          "host transfers of x, f and the output included"}), flush=True)
        return
    if args.kernels_only:
        prob, s = make_solver(T, nels, "mixed")
        names = ALL_KERNELS if args.kernels is None else tuple(args.kernels.split(","))
        kt = kernel_times(T, s, prob, args.reps, args.rmin, names)
        ab = {k: v for k, v in algorithmic_bytes(prob).items() if k in kt}
        heat = None if args.kernels else heat_stats(T, nels, args.reps, args.rel_tol)
        print(json.dumps({"nels": list(nels), "reps": args.reps,
                          "kernel_us": {k: round(v, 3) for k, v in kt.items()},
                          "algorithmic_bytes": ab,
                          "kernel_gbs": {k: round(ab[k] / (kt[k] * 1e-6) / 1e9, 1)
                                         for k in ab},
                          "heat": heat}), flush=True)
        return
    weights = args.weights or default_weights(nels)

    common = (args.steps, args.warmup, args.volfrac, args.rmin, args.rel_tol, args.inner_rel)
    fp64 = run_path(T, "fp64", nels, "fp64", True, *common)
    fp64_mg = run_path(T, "fp64_mg", nels, "fp64_mg", True, *common)
    mixed = run_path(T, "mixed", nels, "mixed", True, *common)
    mixed_mg = run_path(T, "mixed_mg", nels, "mixed_mg", True, *common)
    learned = learned_mg = None
    if weights is not None:
        log(f"warm-start weights: {os.path.relpath(weights, ROOT)}")
        learned = run_path(T, "learned", nels, "mixed", "learned", *common, weights=weights)
        learned_mg = run_path(T, "learned_mg", nels, "mixed_mg", "learned", *common,
                              weights=weights)
    head = learned_mg if learned_mg is not None else mixed_mg
    mma_mg = run_path(T, "mma_mg", nels, "mixed_mg", True, *common, optimizer="mma")

    kt = kernel_times(T, mixed["solver"], mixed["prob"], args.reps, args.rmin)
    ab = algorithmic_bytes(mixed["prob"])
    gbs = {k: ab[k] / (kt[k] * 1e-6) / 1e9 for k in ab}
    log("kernel us: " + ", ".join(f"{k}={v:.2f}" for k, v in kt.items()))
    log("kernel GB/s: " + ", ".join(f"{k}={v:.0f}" for k, v in gbs.items()))
    # the dominant kernel of the headline path: of the launches that run once
    # per PCG iteration at the finest level (or once per V-cycle), the longest
    dom = max(("pcg32", "update32", "mg_pre", "mg_restrict", "mg_coarse", "mg_post"),
              key=lambda k: kt[k])
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_{nels[0]}x{nels[1]}x{nels[2]}.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    comp = assembly_comparator(T, nels, args.volfrac, args.rel_tol)
    log(f"assembled vs matrix-free solve: {comp}")
    heat = heat_stats(T, nels, args.reps, args.rel_tol)
    log(f"heat conduction: {heat}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(nels, args.volfrac, args.rmin, args.rel_tol, args.cpu_seconds)
        log(f"cpu baseline: {cpu['value']:.4f} iter/s ({cpu['sample']}); 1 thread "
            f"{cpu['single_thread']['value']:.5f} iter/s ({cpu['single_thread']['sample']})")

    value = head["iters_per_s"]
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "SIMP iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 / value, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"3D PointLoadCantilever {nels[0]}x{nels[1]}x{nels[2]} hex8, unit "
                         f"load -y at x=max face centre, E0=1 Emin=1e-9 nu=0.3 p=3, volfrac "
                         f"{args.volfrac}, rmin {args.rmin}, OC move 0.2, {args.steps} OC "
                         f"iterations from rho=volfrac; each solve to rel res "
                         f"{args.rel_tol:g}; headline path: FP32 PCG preconditioned by a "
                         f"geometric multigrid V(1,1) cycle ({len(mg_levels(nels))} levels), "
                         f"FP64 residual replacement (every "
                         f"{1 / (args.inner_rel or 1e-3):g}x residual drop)"
                         + (", learned BF16 warm start + previous displacement"
                            if learned_mg is not None else ", previous-displacement warm start")
                         + "; baseline: FP64 Jacobi PCG (the reference's solver)"),
            "ndofs": head["prob"].ndofs, "ncells": head["prob"].ncells,
            "device": devname,
            "weights": None if weights is None else os.path.relpath(weights, ROOT),
        },
        "roofline": {
            "bound": "hbm", "kernel": KERNEL_NAMES[dom],
            "achieved": round(gbs[dom], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs[dom] / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes": ab[dom], "avg_us": round(kt[dom], 3),
            "note": ROOFLINE_NOTES.get(dom),
        },
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in
                                                  ("value", "unit", "cores", "kind", "sample",
                                                   "single_thread")},
        "fp64_path": summary(fp64, fp64),
        "fp64_mg_path": summary(fp64_mg, fp64),
        "mixed_path": summary(mixed, fp64),
        "mixed_mg_path": summary(mixed_mg, fp64),
        "learned_path": None if learned is None else summary(learned, fp64),
        "learned_mg_path": None if learned_mg is None else summary(learned_mg, fp64),
        # the same loop with MMA87 in place of OC (another optimizer, so no
        # trajectory comparison with the OC paths)
        "mma_mg_path": summary(mma_mg, None),
        "speedup_vs_fp64": round(value / fp64["iters_per_s"], 3),
        # the speed-up split into its parts, each a ratio of SIMP iter/s
        "attribution": {
            "multigrid_preconditioner": round(fp64_mg["iters_per_s"] / fp64["iters_per_s"], 3),
            "fp32_arithmetic": round(mixed_mg["iters_per_s"] / fp64_mg["iters_per_s"], 3),
            "fp32_with_jacobi": round(mixed["iters_per_s"] / fp64["iters_per_s"], 3),
            "learned_warm_start": (None if learned_mg is None else
                                   round(learned_mg["iters_per_s"] / mixed_mg["iters_per_s"], 3)),
            "learned_with_jacobi": (None if learned is None else
                                    round(learned["iters_per_s"] / mixed["iters_per_s"], 3)),
        },
        "pcg_solves_per_s": round(value, 4),
        "kernel_us": {k: round(v, 3) for k, v in kt.items()},
        "kernel_gbs": {k: round(v, 1) for k, v in gbs.items()},
        "assembled_vs_matrix_free": comp,
        "heat_conduction": heat,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
