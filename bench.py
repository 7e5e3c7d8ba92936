#!/usr/bin/env python3
"""Benchmark of the dense PMCHWT path on one MI355X.

One step = the whole dense pipeline on device-resident inputs (mesh and basis
uploaded before timing): assemble Z (FP64 + FP32 copies) -> 128 plane-wave
right-hand sides -> FP32 LU on the matrix cores -> multi-RHS solve with FP64
iterative refinement to 1e-10 -> far field, amplitude and Mueller matrices,
orientation average over 64 orientations.

Default workload (the one the metric is quoted on): hexagonal prism, aspect
ratio 0.7, 4 refinements (T = 6144 triangles, N = 9216 RWG, 2N = 18432),
ka = 10, m = 1.3116, 7-point quadrature, Gauss-Legendre 8 x 8 x 1 orientations
x 2 polarisations = 128 RHS, 181 scattering angles.

After the timed steps the same system is factored in FP64 on the GPU (the
FP64-LU comparison and the FP64 reference for the Mueller error), and the FP64
CPU oracle is timed on a bounded sample of the same workload on the host cores.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this benchmark measures
one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "bem-cpp_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

WORKLOADS = {
    # name: (shape, ref, aspect/radius, ka, m, orient grid)
    "c1": ("sphere", 2, 1.0, 1.0, 1.3116 + 0j, (1, 1, 1)),
    "c2": ("sphere", 3, 1.0, 5.0, 1.3116 + 0.01j, (1, 1, 1)),
    "c3": ("hex", 3, 1.4286, 5.0, 1.3116 + 0j, (8, 8, 1)),
    "c5": ("hex", 4, 0.7, 10.0, 1.3116 + 0j, (8, 8, 1)),
}
METRIC = "Z entries/s; s/orient FP32-LU+FP64 IR vs FP64 LU @ Mueller err ≤1e-4"
PEAK_TFLOPS = {"gemm_fp32_3m": 157.3, "gemm_fp64": 78.6}  # MI355X dense MFMA, AMD specification
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


_EXTERNAL_LU = r"""
import json, sys, torch
n = int(sys.argv[1])
torch.backends.cuda.preferred_linalg_library("cusolver")
g = torch.Generator(device="cuda:0")
g.manual_seed(42)
A = torch.randn(n, n, dtype=torch.complex128, device="cuda:0", generator=g)
ms = []
for _ in range(2):
    B = A.clone()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    torch.linalg.lu_factor(B)
    e1.record()
    torch.cuda.synchronize()
    ms.append(e0.elapsed_time(e1))
print(json.dumps({"factor_ms": ms[-1]}))
"""


def external_lu(n):
    """torch.linalg.lu_factor (hipSOLVER -> rocSOLVER zgetrf) on a random
    complex128 n x n matrix: the library FP64 LU that the reference's cuSOLVER
    Zgetrf corresponds to, timed in a child process (torch's HIP runtime and
    this library's cannot share one process).  ms of the second call."""
    import subprocess
    lib = "torch.linalg.lu_factor (hipSOLVER/rocSOLVER zgetrf), complex128, same GPU"
    try:
        r = subprocess.run([sys.executable, "-c", _EXTERNAL_LU, str(n)], capture_output=True, text=True, timeout=240)
        if r.returncode != 0:
            return {"library": lib, "error": (r.stderr or r.stdout)[-300:]}
        out = json.loads(r.stdout.strip().splitlines()[-1])
        return {"library": lib, "n": n, "factor_ms": out["factor_ms"]}
    except Exception as exc:
        return {"library": lib, "error": str(exc)[:300]}


def measured_traffic(workload, kernel_class):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE runs scaled by a copy of
    known size), or None when that workload has not been profiled."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel_class}_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def mueller_err(M, Mref):
    return float((np.abs(M - Mref) / np.abs(Mref[0, 0])[None, None, :]).max())


def lapack_lu_solve(n, nrhs):
    """Seconds of scipy.linalg.lu_factor + lu_solve (LAPACK zgetrf / zgetrs,
    the calls of the reference's CPU twin, python/bem_core.py:1379-1388) on a
    random complex n x n system with nrhs right-hand sides, and the BLAS
    library with its thread count."""
    import scipy.linalg
    lib = "LAPACK zgetrf/zgetrs (scipy.linalg.lu_factor/lu_solve)"
    try:
        from threadpoolctl import threadpool_info
        info = [d for d in threadpool_info() if d.get("user_api") == "blas"]
        if info:
            lib += f" on {info[0].get('internal_api')} {info[0].get('version')}, {info[0].get('num_threads')} threads"
    except Exception:
        pass
    rs = np.random.RandomState(42)
    W = rs.randn(512, 512) + 1j * rs.randn(512, 512)
    scipy.linalg.lu_factor(W, check_finite=False)  # starts the BLAS threads
    Z = rs.randn(n, n) + 1j * rs.randn(n, n)
    B = rs.randn(n, nrhs) + 1j * rs.randn(n, nrhs)
    t0 = time.perf_counter()
    lu = scipy.linalg.lu_factor(Z, overwrite_a=True, check_finite=False)
    scipy.linalg.lu_solve(lu, B, check_finite=False)
    return time.perf_counter() - t0, lib


def cpu_baseline(shape, ref, param, ka, m, n_orient, ntheta, N_full, full=False):
    """The reference's CPU path on the host cores, s/orient of the workload:
    assembly, singular terms, RHS, far field and Mueller by the FP64 oracle
    (OpenMP); the factorisation and the multi-RHS solve by LAPACK, as the
    reference's CPU twin does.  Bounded sample (default): assembly (point
    pairs ~ N^2) on the mesh two refinements coarser, factor + solve (~ n^3)
    of a random system of n = 6144, far field (~ points per orientation) of
    two orientations of the coarser mesh, each scaled to the full size.
    full=True times the whole workload instead (small workloads only)."""
    from oracle import oracle

    threads = oracle.num_threads()
    n_full = 2 * N_full
    # LAPACK first: the oracle's OpenMP threads keep spinning for a while
    # after each parallel region and would compete with the BLAS threads
    n_s = n_full if full else min(6144, n_full)
    t_lu, lapack = lapack_lu_solve(n_s, 2 * n_orient)
    t_lu *= (n_full / n_s) ** 3
    ref_s = ref if full else max(ref - 2, 0)
    p = oracle.Problem.from_shape(shape, ref_s, param)
    scale_pts = (N_full / p.N)
    t0 = time.perf_counter()
    p.assemble_pmchwt(complex(ka), ka * m, 1.0, 1.0 / abs(m))
    t_asm = (time.perf_counter() - t0) * scale_pts ** 2
    n_ff = n_orient if full else 2
    info, Mo, X = p.dense_orient(ka, m, n_ff, 1, 1, ntheta=ntheta)
    t_ff_per_orient = p.last_timings()["farfield_s"] / n_ff * scale_pts
    total = t_asm + t_lu + t_ff_per_orient * n_orient
    if full:
        sample = (f"full workload: FP64 oracle (oracle/bem_oracle.cpp, OpenMP {threads} threads) assembly "
                  f"{t_asm:.3g} s and far field of {n_orient} orientations {t_ff_per_orient * n_orient:.3g} s; "
                  f"{lapack} factor + {2 * n_orient}-RHS solve at n={n_s} {t_lu:.3g} s")
    else:
        sample = (f"FP64 oracle (oracle/bem_oracle.cpp, OpenMP {threads} threads): assembly of {shape} ref {ref_s} "
                  f"(N={p.N}) x (N/{p.N})^2, far field of 2 orientations on ref {ref_s} x points ratio; "
                  f"{lapack} factor + {2 * n_orient}-RHS solve of a random n={n_s} system x (2N/{n_s})^3; "
                  f"extrapolated to N={N_full}: assembly {t_asm:.3g} s, LU+solve {t_lu:.3g} s, far field "
                  f"{t_ff_per_orient * n_orient:.3g} s")
    # near-field corrections of the fast operators on the full mesh (cheap: one
    # same-triangle quadrature per triangle)
    pf = oracle.Problem.from_shape(shape, ref, param)
    t0 = time.perf_counter()
    pf.near_corrections(complex(ka), ka * m)
    t_near = time.perf_counter() - t0
    return {"value": total / n_orient, "unit": "s/orient", "cores": threads, "kind": "port", "sample": sample,
            "near_corrections_ms_full_mesh": t_near * 1e3}


def oracle_fixture_errors(ctx, workload, M32, X32):
    """Errors of this run against the FP64 oracle's fixture of the workload
    (tests/golden/cfg_<workload>.npz, data written by the oracle): Z entry-wise
    on 2000 seeded entries and the diagonal (max |dZ| / max|Z|), surface
    currents (||x - x_ref|| / ||x_ref|| over the stored columns), Mueller
    (max_ij,theta |dM_ij| / M11).  None when there is no fixture."""
    path = os.path.join(ROOT, "tests", "golden", f"cfg_{workload}.npz")
    if not os.path.exists(path):
        return None
    g = np.load(path)
    n = len(g["Zdiag"])
    Ze = ctx.get_Z_entries(g["Zr"], g["Zc"])
    Zd = ctx.get_Z_entries(np.arange(n), np.arange(n))
    zmax = float(g["Zmax"])
    z_err = max(np.abs(Ze - g["Zv"]).max(), np.abs(Zd - g["Zdiag"]).max()) / zmax
    cols = g["Xcols"]
    x_err = max(float(np.linalg.norm(X32[:, c] - g["X"][:, i]) / np.linalg.norm(g["X"][:, i]))
                for i, c in enumerate(cols))
    return {"z_err_vs_oracle": float(z_err), "current_err_vs_oracle": x_err,
            "mueller_err_vs_oracle": mueller_err(M32, g["M"]),
            "fixture": f"tests/golden/cfg_{workload}.npz (FP64 oracle: 2000 seeded Z entries + diagonal, "
                       f"currents of RHS columns {list(int(c) for c in cols)}, Mueller 4x4x{g['M'].shape[2]})"}


def next_rows(ctx, k_ext, k_int, nrhs, ir_tol):
    """The rows beside the dense path, on the same mesh and right-hand sides,
    after the timed steps (each result is its own measurement, not part of
    s/orient): near-field corrections of the fast operators
    (bem_fmm.cu:506-686), GMRES on the FP64 Z preconditioned by the FP32
    factors of Z (block_gmres.h:52-61 with an LU preconditioner) and by the
    block-Jacobi / RAS preconditioner (precond.cu:407-895, 1413-1556)."""
    import bem_amd

    out = {}
    for rep in range(2):  # the first call also builds the Graglia terms of the mesh
        t0 = time.perf_counter()
        rp, ci, vals = ctx.near_corrections(k_ext, k_int)
        ms = (time.perf_counter() - t0) * 1e3
    out["near_corrections"] = {"ms_wall_incl_copy_out": ms, "nnz": int(len(ci)),
                               "entries_per_s": 4 * len(ci) / (ms * 1e-3)}
    log(f"near corrections: {out['near_corrections']}")
    del rp, ci, vals
    info = ctx.factor("fp32")
    if info:
        raise RuntimeError(f"FP32 factorisation info {info}")
    for tol in (1e-5, ir_tol):
        t0 = time.perf_counter()
        _, st = ctx.solve_gmres(None, tol=tol, restart=30, max_iter=200, return_X=False, nrhs=nrhs,
                                allow_noconv=True, precond="lu")
        out[f"gmres_lu32_tol{tol:g}"] = {"iterations": st["iterations"], "converged": bool(st["converged"]),
                                         "residual": st["residual"], "ms": (time.perf_counter() - t0) * 1e3}
        log(f"GMRES, FP32 LU preconditioner, tol {tol:g}: {out[f'gmres_lu32_tol{tol:g}']}")
    # the reference's automatic block-Jacobi parameters for this N and ka
    auto = bem_amd.auto_precond_params(ctx.N, k_ext.real)
    nb, build_ms = ctx.block_jacobi(auto["radius_mult"], auto["max_block_rwg"], auto["overlap_layers"])
    out["block_jacobi"] = {"params": auto, "blocks": nb, "build_ms": build_ms}
    log(f"block-Jacobi build: {out['block_jacobi']}")
    for tol in (1e-5, ir_tol):
        t0 = time.perf_counter()
        _, st = ctx.solve_gmres(None, tol=tol, restart=30, max_iter=300, return_X=False, nrhs=nrhs,
                                allow_noconv=True, precond="block")
        out[f"gmres_block_jacobi_tol{tol:g}"] = {"iterations": st["iterations"], "converged": bool(st["converged"]),
                                                 "residual": st["residual"],
                                                 "ms": (time.perf_counter() - t0) * 1e3}
        log(f"GMRES, block-Jacobi preconditioner, tol {tol:g}: {out[f'gmres_block_jacobi_tol{tol:g}']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS))
    ap.add_argument("--ir-tol", type=float, default=1e-10)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the CPU baseline on the whole workload instead of a bounded sample (small workloads)")
    ap.add_argument("--gemm-tile", default=None,
                    help="fix the FP32 GEMM tile (e.g. 64x64) instead of the trial factorisations; for counter runs, "
                         "whose serialised kernels distort the trials")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="time the steps without the per-launch HIP events (no per-kernel classes or roofline)")
    ap.add_argument("--skip-external", action="store_true",
                    help="skip timing the library FP64 LU (torch.linalg.lu_factor on rocSOLVER)")
    ap.add_argument("--skip-next-rows", action="store_true",
                    help="skip the measurements beside the path (near-field corrections, GMRES, block-Jacobi)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the steps, copy MIB MiB device-to-device with one torch kernel (a known-size "
                         "reference for the FETCH_SIZE / WRITE_SIZE counters of profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path would shard by "
              "orientation: every GPU factors Z and solves its share of the orientations, then one all-reduce "
              "sums the Mueller partial sums)", file=sys.stderr)
        sys.exit(2)

    import bem_amd

    shape, ref, param, ka, m, orient = WORKLOADS[args.workload]
    ntheta = 181
    verts, tris = bem_amd.make_mesh(shape, ref, param)
    RT, w = bem_amd.orientations(*orient) if orient != (1, 1, 1) else bem_amd.lab_single_orientation()
    kh, e0 = bem_amd.planewave_orient_rhs(RT)
    n_orient, nrhs = len(w), len(kh)
    theta = np.pi * np.arange(ntheta) / (ntheta - 1)
    k_ext, k_int, eta_int = complex(ka), ka * m, 1.0 / abs(m)

    ctx = bem_amd.Context(0)
    if args.copy_calibration > 0:
        ctx.device_copy(args.copy_calibration * (1 << 20))
        log(f"copy calibration (k_copy16): {args.copy_calibration} MiB read + {args.copy_calibration} MiB written")
    N = ctx.set_mesh(verts, tris, 7)
    n = 2 * N
    if args.gemm_tile:
        t = ctx.GEMM_TILES.index(args.gemm_tile)
        ctx.gemm_plan("fp32", t, t)
    log(f"workload {args.workload}: {shape} ref {ref} T={len(tris)} N={N} 2N={n} ka={ka} m={m} "
        f"orient={n_orient} rhs={nrhs}")

    def step():
        ctx.assemble(k_ext, k_int, 1.0, eta_int, z64=True, z32=True)
        ta = ctx.timings()["assembly_ms"]
        ctx.rhs_planewave(k_ext, 1.0, kh, e0)
        tr = ctx.timings()["rhs_ms"]
        info = ctx.factor("fp32")
        if info:
            raise RuntimeError(f"FP32 factorisation info {info}")
        _, st = ctx.solve(None, ir_tol=args.ir_tol, ir_max=20, return_X=False, nrhs=nrhs, allow_noconv=True)
        t = ctx.timings()
        M = ctx.orient_mueller(k_ext, 1.0, RT, w, theta)
        t = ctx.timings()
        return M, st, dict(assembly=ta, rhs=tr, factor=t["factor_ms"], solve=t["solve_ms"], refine=t["refine_ms"],
                           farfield=t["farfield_ms"])

    for i in range(args.warmup):
        t0 = time.perf_counter()
        step()
        log(f"warmup {i}: {time.perf_counter() - t0:.3f} s")
    # the timed steps carry HIP events only on the dominant kernel (the LU
    # trailing-update GEMM): an event costs the device ~2 us, so events on all
    # ~3700 launches of a factorisation would slow it by ~10 %
    if not args.no_kernel_timing:
        ctx.kernel_stats("update")
    phases = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        M32, st, ph = step()
        phases.append(ph)
        log(f"step {i}: {time.perf_counter() - ts:.3f} s  {ph}  IR sweeps {st['iterations']} "
            f"residual {st['residual']:.2e}")
    elapsed = time.perf_counter() - t0
    kstats = ctx.kernel_stats()
    ctx.kernel_stats("off")
    ms_per_step = elapsed / args.steps * 1e3
    # per-class breakdown from one more step with events on every launch
    # (outside the timed steps; its times include the events' cost)
    kstats_all = {}
    if not args.no_kernel_timing:
        ctx.kernel_stats("on")
        _, _, ph_instr = step()
        kstats_all = ctx.kernel_stats()
        ctx.kernel_stats("off")
        log(f"instrumented step: {ph_instr}")
    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}

    # adjoint gradient of the averaged-M11 loss L = sum_t w_t log M11(theta_t)
    # (w uniform in cos theta) with respect to (Re m, Im m, ka), through the
    # FP32 factorisation of the last step (one adjoint solve + dZ/dp contraction)
    wt = np.sin(theta) * np.pi / (len(theta) - 1)
    G = np.zeros((4, 4, len(theta)))
    G[0, 0] = wt / M32[0, 0]
    t_g = time.perf_counter()
    grad = ctx.mueller_grad(G)
    adjoint = dict(loss="sum_t w_t log M11(theta_t), w uniform in cos theta",
                   grad_re_m_im_m_ka=[float(x) for x in grad], device_ms=float(ctx.timings()["adjoint_ms"]),
                   wall_ms=(time.perf_counter() - t_g) * 1e3)
    log(f"adjoint gradient: {adjoint}")

    # FP64 LU on the same GPU: the comparison and the FP64 reference for the Mueller error
    plan32 = ctx.gemm_plan_info("fp32")
    fp64 = {}
    plan64 = None
    # the FP32 + refinement solutions of the last step against the oracle's fixture
    X32, _ = ctx.solve(None, ir_tol=args.ir_tol, ir_max=20, return_X=True, nrhs=nrhs, allow_noconv=True)
    vs_oracle = oracle_fixture_errors(ctx, args.workload, M32, X32)
    del X32
    log(f"against the oracle fixture: {vs_oracle}")
    if not args.skip_fp64:
        ctx.factor("fp64")  # first FP64 factorisation of this size: tile trials + factorisation (untimed)
        plan64 = ctx.gemm_plan_info("fp64")
        f64_runs = []
        for _ in range(3):
            ctx.factor("fp64")
            f64_runs.append(ctx.timings()["factor_ms"])
        f64_factor = float(np.mean(f64_runs))
        X, st64 = ctx.solve(None, ir_tol=0.0, ir_max=0, return_X=False, nrhs=nrhs)
        f64_solve = ctx.timings()["solve_ms"]
        M64 = ctx.orient_mueller(k_ext, 1.0, RT, w, theta)
        f64_ff = ctx.timings()["farfield_ms"]
        fp64 = dict(factor_ms=f64_factor, factor_ms_runs=f64_runs, solve_ms=f64_solve, farfield_ms=f64_ff,
                    s_per_orient=(ph["rhs"] + f64_factor + f64_solve + f64_ff) / n_orient / 1e3,
                    mueller_err_fp32ir_vs_fp64=mueller_err(M32, M64))
        log(f"fp64: {fp64}")

    rows = None
    if not args.skip_next_rows:
        rows = next_rows(ctx, k_ext, k_int, nrhs, args.ir_tol)
        log(f"next rows: {rows}")

    # GEMM tile plan picked by the autotuner at this size (timed again here, outside the steps)
    tune = ctx.autotune(nrhs)

    # library FP64 LU of the same size on the same GPU (external reference)
    external = None
    if not args.skip_external:
        external = external_lu(n)
        log(f"external FP64 LU: {external}")

    # roofline of the dominant kernel, the FP32 3M trailing-update GEMM of the
    # LU (k_cupdate_nt), from the live HIP-event times of its launches over the
    # timed steps
    dom = "gemm_fp32_update"
    launches, kms, kflops, kbytes = kstats.get(dom, (0, 0.0, 0.0, 0.0))
    if launches > 0 and kms > 0:
        ach = kflops / (kms * 1e-3) / 1e12
        tile = plan32["update"] if plan32 else "?"
        roof = {"kernel": f"k_cupdate_nt (3M complex GEMM C -= A Bt^T, both operands k-contiguous, LU trailing "
                          f"updates, {tile} tile; k_cgemm<float, cfg, 1> for the last updates below 64 rows)",
                "bound": "mfma",
                "achieved": ach, "peak": PEAK_TFLOPS["gemm_fp32_3m"], "unit": "TFLOP/s",
                "frac": ach / PEAK_TFLOPS["gemm_fp32_3m"], "traffic": measured_traffic(args.workload, dom),
                "launches": int(launches), "avg_launch_ms": kms / launches,
                "flops_per_launch": kflops / launches}
    else:
        roof = None

    step_s = ms_per_step / 1e3
    s_orient_excl_asm = (ph["rhs"] + ph["factor"] + ph["solve"] + ph["refine"] + ph["farfield"]) / n_orient / 1e3
    out = {
        "metric": METRIC,
        "value": step_s / n_orient,
        "unit": "s/orient",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "precision": "fp32 LU/solve on MFMA (3M) + fp64 refinement; fp64 assembly sums, far field, Mueller",
        "data": "synthetic (deterministic mesh and plane waves)",
        "config": {
            "workload": (f"{shape} {'AR ' + str(param) if shape == 'hex' else 'r=1'} ref {ref}: T={len(tris)}, "
                         f"N={N}, 2N={n}; ka={ka}, m={m.real}{m.imag:+}i; 7-point quadrature; "
                         f"{n_orient} orientations x 2 polarisations = {nrhs} RHS; {ntheta} angles; "
                         "adjoint gradient of the averaged-M11 loss reported beside the timed step"),
            "n": n, "nrhs": nrhs, "n_orient": n_orient, "ntheta": ntheta,
        },
        "z_entries_per_s": n * n / (ph["assembly"] * 1e-3),
        "gemm_plan": {"fp32": plan32, "fp64": plan64,
                      "fp32_trailing_gemm_alone_tflops": tune["fp32_trailing"],
                      "fp64_trailing_gemm_alone_tflops": tune["fp64_trailing"]},
        "external_fp64_lu": external,
        "s_per_orient_fp32ir_excl_assembly": s_orient_excl_asm,
        "s_per_orient_fp64_excl_assembly": fp64.get("s_per_orient"),
        "speedup_fp32ir_vs_fp64": (fp64["s_per_orient"] / s_orient_excl_asm) if fp64 else None,
        "mueller_err_vs_fp64": fp64.get("mueller_err_fp32ir_vs_fp64"),
        "vs_oracle": vs_oracle,
        "ir_sweeps": st["iterations"], "ir_residual": st["residual"],
        "phase_ms": ph,
        "fp64_phase_ms": {k: fp64[k] for k in ("factor_ms", "factor_ms_runs", "solve_ms", "farfield_ms")}
        if fp64 else None,
        "kernels_instrumented_step": {
            k: {"launches": int(v[0]), "ms": v[1], "tflops": (v[2] / (v[1] * 1e-3) / 1e12) if v[1] else 0.0}
            for k, v in kstats_all.items() if v[0] > 0},
        "adjoint_gradient": adjoint,
        "beside_the_path": rows,
        "roofline": roof,
    }
    ctx.close()
    if not args.skip_cpu:
        out["cpu_baseline"] = cpu_baseline(shape, ref, param, ka, m, n_orient, ntheta, N, full=args.cpu_full)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
