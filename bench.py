"""Headline benchmark: assembly + CG DOFs/s on one MI355X.

Workload (default): 3-D Poisson -div(k grad u) = f on the unit cube, 128^3
hexahedral Q1 cells (2,048,383 free DOFs), k = 1 + 0.5 sin(2pi x) sin(2pi y)
sin(2pi z), f manufactured from u = sin(pi x) sin(pi y) sin(pi z), homogeneous
Dirichlet, Jacobi-PCG to 1e-10 relative residual (RES_RES).

One step = integrate + assemble on the device + solve from x0 = 0 (inputs,
DOF lists and the CSR pattern already resident in HBM).  value = free DOFs /
(assembly + CG) time of the plain FP64 path.  The FP32-refined path runs the
same steps and is reported beside it (speed-up and error vs the FP64 solution).
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

HBM_PEAK_GBS = 8000.0


def make_problem(F, env, n, order=1):
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": 3, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": [n, n, n],
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": [0, 1, 0, 1, 0, 1]})
    fes = F.serial_fe_space_t().create(tria, order=order)
    di = F.poisson_discrete_integration_t(k="sin3", f="manufactured")
    op = F.fe_affine_operator_t()
    op.create("CSR", [True], fes, di)
    return tria, fes, di, op


def make_solver(F, env, op, precision, rtol, delta=0.0):
    s = F.iterative_linear_solver_t().create(env)
    s.set_type_from_string("CG")
    s.set_parameters_from_pl({"ILS_RTOL": rtol, "ILS_MAX_NUM_ITERATIONS": 20000, "FEMHIP_PRECISION": precision,
                              "FEMHIP_INNER_RTOL": delta})
    s.set_operators(op.get_tangent(), F.jacobi_preconditioner_t(op))
    return s


def run_steps(op, solver, steps):
    t_asm, t_cg, wall, last = [], [], [], None
    for _ in range(steps):
        t0 = time.perf_counter()
        op.compute()                       # returns after the device finished (event sync)
        r = solver.solve_resident()        # returns after the device finished
        wall.append(time.perf_counter() - t0)
        t_asm.append(op.last_compute_ms / 1e3)
        t_cg.append(r["solve_ms"] / 1e3)
        last = r
    return np.array(t_asm), np.array(t_cg), np.array(wall), last


def grf_k_cell(n, sigma=1.5, ell=0.1, seed=20261202):
    """k per cell = exp(mean of g over the cell's 8 vertices), g a Gaussian
    random field on the (n+1)^3 vertex grid of the unit cube: white noise from
    default_rng(seed) filtered in Fourier space by the square root of the
    squared-exponential spectrum (correlation length ell), normalised to mean 0
    and standard deviation sigma.  Array axes are (z, y, x), so the flattened
    cell array is x-fastest like the mesh's cells."""
    m = n + 1
    w = np.random.default_rng(seed).standard_normal((m, m, m))
    kf = 2 * np.pi * np.fft.fftfreq(m, d=1.0 / n)
    kr = 2 * np.pi * np.fft.rfftfreq(m, d=1.0 / n)
    k2 = kf[:, None, None] ** 2 + kf[None, :, None] ** 2 + kr[None, None, :] ** 2
    g = np.fft.irfftn(np.fft.rfftn(w) * np.exp(-(ell ** 2) * k2 / 4.0), s=w.shape, axes=(0, 1, 2))
    g = (g - g.mean()) / g.std() * sigma
    c = sum(g[a:m - 1 + a, b:m - 1 + b, d:m - 1 + d] for a in (0, 1) for b in (0, 1) for d in (0, 1)) / 8.0
    return np.exp(c).ravel()


def measure_config(F, env, n, order, di, rtol, steps, warmup, with_mixed=True, exact=None, strict=False):
    """Assembly + Jacobi-PCG DOFs/s of one configuration in FP64 and FP32-refined
    (and FP32_STRICT: the FP32 matrix integrated and summed in FP32)."""
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": 3, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": [n, n, n],
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": [0, 1, 0, 1, 0, 1]})
    t0 = time.perf_counter()
    fes = F.serial_fe_space_t().create(tria, order=order)
    op = F.fe_affine_operator_t()
    op.create("CSR", [True], fes, di)
    out = {"n_free_dofs": fes.num_free_dofs, "nnz_full": op.nnz_full, "setup_s": time.perf_counter() - t0}
    xs, solvers = {}, []
    precs = ("FP64", "FP32_REFINED") if with_mixed else ("FP64",)
    for prec in precs + (("FP32_STRICT",) if with_mixed and strict else ()):
        s = make_solver(F, env, op, prec, rtol)
        solvers.append(s)
        first = None
        for _ in range(warmup):   # the first step also lays out the solver's matrix copy
            t1 = time.perf_counter()
            run_steps(op, s, 1)
            first = time.perf_counter() - t1 if first is None else first
        ta, tc, wall, r = run_steps(op, s, steps)
        xs[prec] = s.get_solution()
        out[prec] = {"value": fes.num_free_dofs / float(np.mean(ta + tc)), "t_assemble_ms": float(np.mean(ta) * 1e3),
                     "t_cg_ms": float(np.mean(tc) * 1e3), "iterations": r["num_iterations"],
                     "outer_iterations": r["outer_iterations"], "true_rel_residual": r["true_residual"] / r["rhs_norm"],
                     "first_step_s": first}
        if exact:
            out[prec]["error_vs_u"] = s.error_norms(exact)
        # the CG SpMV of this configuration back to back (HIP events, 50 launches)
        sp_ms, sp_bytes = s.profile_spmv(50)
        out[prec]["spmv"] = {"kernel_ms": sp_ms, "algorithmic_bytes": sp_bytes, "GBs": sp_bytes / sp_ms / 1e6,
                             "frac_of_hbm_peak": sp_bytes / sp_ms / 1e6 / HBM_PEAK_GBS}
    if with_mixed:
        out["speedup_mixed_vs_fp64"] = out["FP32_REFINED"]["value"] / out["FP64"]["value"]
        out["rel_err_mixed_vs_fp64"] = float(np.linalg.norm(xs["FP32_REFINED"] - xs["FP64"]) /
                                             np.linalg.norm(xs["FP64"]))
    if "FP32_STRICT" in out:
        out["speedup_strict_vs_fp64"] = out["FP32_STRICT"]["value"] / out["FP64"]["value"]
        out["rel_err_strict_vs_fp64"] = float(np.linalg.norm(xs["FP32_STRICT"] - xs["FP64"]) /
                                              np.linalg.norm(xs["FP64"]))
    for s in solvers:
        s.free()
    op.free()
    fes.free()
    return out


def cpu_baseline(n, iters_full, op_matrix=None, rhs=None):
    """Oracle (CPU restatement of the reference algorithms), 1 thread: assembly
    of a z-slab of 1/8 of the cells scaled x8, plus 10 Jacobi-PCG iterations on
    the full matrix scaled to the full solve's iteration count."""
    from oracle import oracle as O
    sp = O.Space(O.make_problem(3, (n, n, n), 1))
    ph = O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_MANUFACTURED)
    slab = sp.num_cells // 8
    t0 = time.perf_counter()
    sp.assemble(ph, True, cells=(0, slab))
    t_asm = (time.perf_counter() - t0) * (sp.num_cells / slab)
    irp, ja, val = op_matrix.irp, op_matrix.ja, op_matrix.val
    k = 10
    t0 = time.perf_counter()
    O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-30, maxit=k)
    t_it = (time.perf_counter() - t0) / k
    total = t_asm + t_it * iters_full
    out = {"value": sp.n_free / total, "unit": "DOFs/s", "cores": 1, "kind": "port",
           "sample": f"oracle assembly of cells [0,{slab}) of the {n}^3 mesh (x8) + {k} Jacobi-PCG iterations "
                     f"on the full matrix (x{iters_full} iterations); t_asm={t_asm:.2f}s t_iter={t_it*1e3:.1f}ms",
           "includes": "COO append + sort/compress of the assembly (the GPU value starts from a resident pattern)"}
    out["all_cores"] = cpu_baseline_all_cores(sp, ph, irp, ja, val, rhs, iters_full)
    return out


def cpu_baseline_all_cores(sp, ph, irp, ja, val, rhs, iters_full):
    """All host threads (not the reference's serial algorithm): the oracle's
    assembly of one z-slab per thread run concurrently (ctypes calls release
    the GIL; the slabs' COO merge is not timed), plus 10 iterations of an
    OpenMP Jacobi-PCG on the full-storage matrix, scaled to the solve."""
    import concurrent.futures as cf
    import scipy.sparse as sps
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    cells = sp.num_cells
    bounds = [(cells * i // cores, cells * (i + 1) // cores) for i in range(cores)]
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(lambda c: sp.assemble(ph, True, cells=c), bounds))
    t_asm = time.perf_counter() - t0
    n = len(irp) - 1
    U = sps.csr_matrix((val, ja - 1, irp - 1), shape=(n, n))
    A = (U + U.T - sps.diags(U.diagonal())).tocsr()
    A.sort_indices()
    firp, fja = (A.indptr + 1).astype(np.int32), (A.indices + 1).astype(np.int32)
    k = 10
    t0 = time.perf_counter()
    O.cg_threads(firp, fja, A.data, rhs, rtol=1e-30, maxit=k)
    t_it = (time.perf_counter() - t0) / k
    total = t_asm + t_it * iters_full
    return {"value": sp.n_free / total, "unit": "DOFs/s", "cores": cores,
            "kind": "port (threaded slab assembly + OpenMP CG; not the reference's serial algorithm)",
            "sample": f"{cores} concurrent slab assemblies covering all cells + {k} OpenMP Jacobi-PCG iterations "
                      f"(x{iters_full}); t_asm={t_asm:.2f}s t_iter={t_it*1e3:.1f}ms"}


def batched_problems(count=4096, seed=4096):
    """Batched workload: 70 % 2-D Q1 n x n with n log-uniform in [16,256], 30 % 3-D
    Q1 n^3 with n log-uniform in [8,48]; constant k uniform in [0.1,10]."""
    rng = np.random.default_rng(seed)
    is2d = rng.random(count) < 0.7
    n2 = np.rint(np.exp(rng.uniform(np.log(16), np.log(256), count))).astype(int)
    n3 = np.rint(np.exp(rng.uniform(np.log(8), np.log(48), count))).astype(int)
    k = rng.uniform(0.1, 10.0, count)
    return [(2, int(n2[i]), float(k[i]), 1.0) if is2d[i] else (3, int(n3[i]), float(k[i]), 1.0)
            for i in range(count)]


def run_batched(F, env, steps, rtol=1e-8):
    problems = batched_problems()
    b = F.batch_t(env, problems, rtol=rtol)
    t_asm = b.assemble()
    threads = b.autotune()
    b.solve()   # warm-up (graph instantiation)
    runs = [b.solve() for _ in range(max(1, steps))]
    ms = np.array([r["ms"] for r in runs])
    last = runs[-1]
    lat = last["latency_us"] / 1e3
    dofs = int(np.sum(last["num_dofs"]))
    rel = last["true_residual"] / np.maximum(last["rhs_norm"], 1e-300)
    out = {"problems": len(problems), "total_free_dofs": dofs, "assemble_ms": t_asm,
           "solve_ms": float(ms.mean()), "solves_per_s": len(problems) / (ms.mean() / 1e3),
           "dofs_per_s": dofs / ((t_asm + ms.mean()) / 1e3),
           "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
           "latency_ms_max": float(lat.max()), "converged": int(np.sum(last["did_converge"])),
           "max_true_rel_residual": float(rel.max()), "threads_per_block": threads,
           "iterations_total": int(np.sum(last["num_iterations"]))}
    # roofline of the solve launch: FP32 inner iterations of every problem,
    # (4 3^d + 40) bytes per row and iteration (the diagonals once, ten 4-byte
    # vector passes), over the launch time
    dims = np.array([p[0] for p in problems])
    per_row = 4.0 * 3.0 ** dims + 40.0
    alg = float(np.sum(np.asarray(last["num_iterations"], dtype=float) * np.asarray(last["num_dofs"], dtype=float)
                       * per_row))
    gbs = alg / (ms.mean() / 1e3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": alg,
                       "definition": "sum over problems of inner iterations x free DOFs x (4*3^d + 40) B / solve_ms"}
    b.free()
    return out


def run_multifield(F, env, n=48, rtol=1e-10):
    """Two coupled fields on the unit cube, n^3 cells: u in Q2 and v in Q1
    (the shapes of a velocity-pressure pair), each fixed on the boundary, in
    two blocks; kappa = (1, 1), sigma = 1, f = (1, 1).  Block assembly on the
    device, block Jacobi-PCG (RES_RES, rtol) over both blocks."""
    lim = [0, 1, 0, 1, 0, 1]
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": 3, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": [n, n, n],
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": lim})
    t0 = time.perf_counter()
    spaces = [F.serial_fe_space_t().create(tria, order=2), F.serial_fe_space_t().create(tria, order=1)]
    layout = F.block_layout_t().create(2, [1, 2])
    op = F.block_fe_affine_operator_t()
    op.create("CSR", [False], spaces, layout, F.coupled_diffusion_discrete_integration_t((1.0, 1.0), 1.0, (1.0, 1.0)))
    setup_s = time.perf_counter() - t0
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": rtol, "ILS_MAX_NUM_ITERATIONS": 20000})
    solver.set_operators(op, F.jacobi_preconditioner_t(op))
    op.compute()
    solver.solve()                      # warm-up
    asm, sol = [], []
    for _ in range(2):
        op.compute()
        asm.append(op.last_compute_ms)
        solver.solve()
        sol.append(solver.result["solve_ms"])
    r = solver.result
    nnz = int(np.sum(op.block_nnz))
    ndofs = op.num_rows
    it = r["num_iterations"]
    # per iteration: every block's CSR (8 B value + 4 B column per entry, row
    # pointers) once in the SpMV, and about 13 vector passes of 8 B per row
    per_it = 12.0 * nnz + 4.0 * (ndofs + 2) + 104.0 * ndofs
    out = {"workload": f"two fields on {n}^3 cells: Q2 + Q1, two blocks", "dofs": ndofs,
           "block_sizes": op.block_sizes, "nnz": nnz, "setup_s": setup_s,
           "t_assemble_ms": float(np.mean(asm)), "t_cg_ms": float(np.mean(sol)), "iterations": it,
           "true_rel_residual": r["true_residual"] / r["rhs_norm"],
           "value": ndofs / ((np.mean(asm) + np.mean(sol)) / 1e3), "unit": "DOFs/s",
           "ms_per_iteration": float(np.mean(sol)) / max(1, it),
           "cg_GBs": per_it * it / (np.mean(sol) / 1e3) / 1e9}
    out["cg_frac_of_hbm_peak"] = out["cg_GBs"] / HBM_PEAK_GBS
    solver.free()
    op.free()
    for sp in spaces:
        sp.free()
    return out


SERVE_RATES = (10, 20, 30, 100, 200, 500, 1000, 2000, 5000)   # below and above the ~30 solves/s capacity
SERVE_POLICY_RATES = (30, 1000)                                   # also served smallest-first


def serving_requests(count, seed=5000):
    """Serving workload: vector-Q1 elasticity on n^3 with n log-uniform in [8,96],
    E uniform in [1,200], nu uniform in [0.2,0.45], clamped on x = 0, body force
    (0,0,-1)."""
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(np.log(8), np.log(96), count))).astype(int)
    E = rng.uniform(1.0, 200.0, count)
    nu = rng.uniform(0.2, 0.45, count)
    return [(int(n[i]), float(E[i]), float(nu[i])) for i in range(count)]


def run_serving(F, env, count, rates=SERVE_RATES, rtol=1e-8, policy_rates=SERVE_POLICY_RATES):
    """Open-loop Poisson arrivals at each rate: the same `count` requests are
    submitted with arrival times from exponential gaps (seed 5000 + rate); the
    server admits them as slots and HBM budget free up, first come first
    served ("rates") and, at policy_rates, smallest estimate first with a 2 s
    aging bound ("smallest_first").  Latency = arrival to solution on the
    host.  Roofline: FP32 inner iterations x DOFs x 364 B (81 FP32 stencil
    values and ten 4-byte vector passes per row) over the span."""
    reqs = serving_requests(count)
    out = {"requests_per_rate": count, "rates": {}, "smallest_first": {}}
    for lam, policy in [(r, "fcfs") for r in rates] + [(r, "smallest") for r in policy_rates]:
        srv = F.serve_t(env, hbm_budget_bytes=96 << 30, max_active=8, setup_threads=8, rtol=rtol,
                        admission=policy, aging_s=2.0 if policy == "smallest" else 0.0)
        gaps = np.random.default_rng(5000 + lam).exponential(1.0 / lam, count)
        t0 = srv.now() + 0.02
        arrivals = t0 + np.cumsum(gaps) - gaps[0]
        tickets = [srv.submit(n, E, nu, arrival_s=float(a)) for (n, E, nu), a in zip(reqs, arrivals)]
        res = [srv.wait(t, timeout_s=600) for t in tickets]
        end = srv.now()
        st = srv.stats()
        srv.free()
        lat = np.array([r["latency_ms"] for r in res])
        dofs = sum(r["num_dofs"] for r in res)
        span = end - arrivals[0]
        alg = float(sum(r["num_iterations"] * r["num_dofs"] for r in res)) * 364.0
        out["rates" if policy == "fcfs" else "smallest_first"][str(lam)] = {
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "hbm_GBs": alg / span / 1e9, "hbm_frac": alg / span / 1e9 / HBM_PEAK_GBS,
            "max_ms": float(lat.max()), "solves_per_s": count / span, "dofs_per_s": dofs / span,
            "mean_queue_ms": float(np.mean([r["queue_ms"] for r in res])),
            "mean_setup_ms": float(np.mean([r["setup_ms"] for r in res])),
            "mean_solve_ms": float(np.mean([r["solve_ms"] for r in res])),
            "converged": int(sum(r["did_converge"] for r in res)),
            "max_true_rel_residual": float(max(r["true_residual"] / max(r["rhs_norm"], 1e-300) for r in res)),
            "peak_active": st["peak_active"], "running_peak": st["running_peak"],
            "hbm_peak_GB": st["hbm_peak"] / 1e9}
        print(f"[serve] rate {lam}/s {policy}: {out['rates' if policy == 'fcfs' else 'smallest_first'][str(lam)]}",
              file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--order", type=int, default=1, help="Lagrangian order of the --spmv-only problem")
    ap.add_argument("--delta", type=float, default=0.0,
                    help="reliable-update factor of the FP32-refined solver (0: the library default 0.1)")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-multifield", action="store_true")
    ap.add_argument("--multifield-only", action="store_true", help="only the two-field block configuration")
    ap.add_argument("--no-q2", action="store_true", help="skip the 64^3 Q2 configuration")
    ap.add_argument("--no-grf", action="store_true", help="skip the 192^3 random-field mixed-precision configuration")
    ap.add_argument("--configs-only", action="store_true", help="only the Q2 and random-field configurations")
    ap.add_argument("--serve-requests", type=int, default=200, help="requests per arrival rate")
    ap.add_argument("--serving-only", action="store_true", help="only the serving sweep (printed as JSON)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only launch the FP64 and FP32 SpMV kernels 50 times each (for rocprofv3 --pmc runs)")
    ap.add_argument("--assembly-only", action="store_true",
                    help="only assemble the --n^3 problem of --order 10 times (Q1: with the vector kernel, then "
                         "10 times with the FP64 matrix-core kernel; Q2: element matrices on the matrix cores) "
                         "(for rocprofv3 --pmc runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device-to-device 5 times (counter calibration for --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the path would shard by subdomains with one interface exchange, which is not built.",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        src = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 8, "copies": 5}))
        return

    import fempar_amd as F
    env = F.environment_t(0)
    n = args.n
    if args.serving_only:
        print(json.dumps(run_serving(F, env, args.serve_requests)))
        return

    def other_configs():
        res = {}
        if not args.no_q2:
            di = F.poisson_discrete_integration_t(k="sin3", f="manufactured")
            res["poisson3d_q2_64"] = measure_config(F, env, 64, 2, di, args.rtol, args.steps, args.warmup,
                                                    exact="sin")
            print(f"[bench] q2 64^3: {res['poisson3d_q2_64']}", file=sys.stderr, flush=True)
        if not args.no_grf:
            t0 = time.perf_counter()
            kc = grf_k_cell(192)
            di = F.poisson_discrete_integration_t(k="cell", k_cell=kc, f="const", f0=1.0)
            r = measure_config(F, env, 192, 1, di, args.rtol, 1, 1, strict=True)
            r["k_contrast"] = float(kc.max() / kc.min())
            r["field_s"] = time.perf_counter() - t0 - r["setup_s"]
            res["poisson3d_q1_192_grf"] = r
            print(f"[bench] grf 192^3: {r}", file=sys.stderr, flush=True)
        return res

    if args.configs_only:
        print(json.dumps(other_configs()))
        return
    if args.assembly_only:
        tria, fes, di, op = make_problem(F, env, n, order=args.order)
        op.compute()                       # builds the owner-computes maps once
        res = {}
        for variant, flag in ((("vector", "0"), ("mfma", "1")) if args.order == 1 else (("mfma", "0"),)):
            os.environ["FEMHIP_ASSEMBLY_MFMA"] = flag
            ms = []
            for _ in range(10):
                op.compute()
                ms.append(op.last_compute_ms)
            res[variant] = {"compute_ms_median": float(np.median(ms))}
        os.environ.pop("FEMHIP_ASSEMBLY_MFMA")
        print(json.dumps(res))
        return
    if args.multifield_only:
        print(json.dumps(run_multifield(F, env)))
        return
    if args.spmv_only:
        tria, fes, di, op = make_problem(F, env, n, order=args.order)
        op.compute()
        res = {}
        for prec in ("FP64", "FP32_REFINED"):
            ms, by = make_solver(F, env, op, prec, args.rtol).profile_spmv(50)
            res[prec] = {"kernel_ms": ms, "algorithmic_bytes": by, "GBs": by / ms / 1e6}
        print(json.dumps(res))
        return
    t0 = time.perf_counter()
    tria, fes, di, op = make_problem(F, env, n)
    t_setup = time.perf_counter() - t0
    s64 = make_solver(F, env, op, "FP64", args.rtol)
    t1 = time.perf_counter()
    for _ in range(args.warmup):   # the first step also lays out the solver's matrix copy
        run_steps(op, s64, 1)
    t_first = (time.perf_counter() - t1) / max(1, args.warmup)
    ta, tc, wall, r64 = run_steps(op, s64, args.steps)
    x64 = s64.get_solution()
    err64 = s64.error_norms("sin")   # vs the manufactured u = sin sin sin
    ms_step = float(np.mean(wall) * 1e3)
    value = fes.num_free_dofs / float(np.mean(ta + tc))

    # roofline of the dominant kernel (CG SpMV): HIP events around 50
    # back-to-back launches on the solver stream (agrees with the kernel trace);
    # beside it the event mean over the SpMVs of one solve, whose start events
    # also hold each launch's dispatch latency
    spmv_ms, spmv_bytes = s64.profile_spmv(50)
    insolve_ms, insolve_its, _ = s64.profile_cg()
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"spmv_traffic_poisson3d_q1_{n}.json")
    if os.path.exists(tpath):   # HBM bytes per launch from the committed rocprofv3 --pmc runs
        traffic = json.load(open(tpath)).get("FP64", {}).get("hbm_bytes_per_launch")

    def run_mixed(precision):
        s32 = make_solver(F, env, op, precision, args.rtol, args.delta)
        for _ in range(args.warmup):
            run_steps(op, s32, 1)
        ta32, tc32, wall32, r32 = run_steps(op, s32, args.steps)
        x32 = s32.get_solution()
        sp32_ms, sp32_bytes = s32.profile_spmv(50)
        res = {
            "precision": precision,
            "value": fes.num_free_dofs / float(np.mean(ta32 + tc32)),
            "ms_per_step": float(np.mean(wall32) * 1e3),
            "t_assemble_ms": float(np.mean(ta32) * 1e3), "t_cg_ms": float(np.mean(tc32) * 1e3),
            "inner_iterations": r32["num_iterations"], "outer_iterations": r32["outer_iterations"],
            "speedup_vs_fp64": float(np.mean(ta + tc) / np.mean(ta32 + tc32)),
            "rel_err_vs_fp64": float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)),
            "true_rel_residual": r32["true_residual"] / r32["rhs_norm"],
            "spmv_fp32_GBs": sp32_bytes / (sp32_ms * 1e-3) / 1e9,
            "error_vs_u": s32.error_norms("sin"),
        }
        s32.free()
        return res

    mixed = None if args.no_mixed else run_mixed("FP32_REFINED")
    # the FP32 matrix integrated in FP32 on the matrix cores and summed in FP32,
    # refined against the same FP64 residual
    mixed_strict = None if args.no_mixed else run_mixed("FP32_STRICT")

    batched = None if args.no_batched else run_batched(F, env, args.steps)
    serving = None if args.no_serving else run_serving(F, env, args.serve_requests)
    configs = other_configs()
    if not args.no_multifield and not args.configs_only:
        configs["multifield_q2q1_48"] = run_multifield(F, env)
        print(f"[bench] multifield: {configs['multifield_q2q1_48']}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        A = op.get_matrix()
        cpu = cpu_baseline(n, r64["num_iterations"], A, op.get_translation())

    out = {
        "metric": "assembly+CG DOFs/s, Q1 128^3 Poisson, mixed vs FP64; batched p99 solve latency",
        "value": value, "unit": "DOFs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"poisson3d_q1_{n}^3_sin3k_jacobi_pcg_rtol{args.rtol:g}",
                   "n_free_dofs": fes.num_free_dofs, "nnz_full": op.nnz_full, "nnz_stored_upper": op.nnz,
                   "cells": fes.num_cells},
        "t_assemble_ms": float(np.mean(ta) * 1e3), "t_cg_ms": float(np.mean(tc) * 1e3),
        "cg_iterations": r64["num_iterations"], "true_rel_residual": r64["true_residual"] / r64["rhs_norm"],
        "error_vs_u": err64,
        "setup_s": t_setup, "warmup_step_s": t_first,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": (f"profiles/spmv_traffic_poisson3d_q1_{n}.json: rocprofv3 --pmc FETCH_SIZE and "
                                        "WRITE_SIZE of this kernel at this size, calibrated by a plain copy; "
                                        "read from the committed file, not measured in this run")
                     if traffic is not None else None,
                     "kernel": "k_spmv<double> (FP64 CG SpMV on the sliced solver copy)",
                     "kernel_ms": spmv_ms, "algorithmic_bytes": spmv_bytes,
                     "kernel_ms_in_solve_events": insolve_ms, "in_solve_launches": insolve_its},
        "mixed": mixed,
        "mixed_strict": mixed_strict,
        "batched": batched,
        "serving": serving,
        "configs": configs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
