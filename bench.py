#!/usr/bin/env python3
"""Benchmark of the MI355X shell-assembly / buckling path.

Workload (the line's `value`): assembly of the stiffness matrix K of the
200 x 400 axially compressed cylinder (R = 1, L = 2, 80,000 4-node shell
elements, 80,400 nodes, 482,400 dofs) with seeded per-element thickness
U[0.0015, 0.0025] (seed 1234) and a seeded radial imperfection of amplitude
0.1 * 0.002 (wavenumbers from seed 1235); E = 70 GPa, nu = 0.3.  One step =
every element Jacobian of the linear-model shell, scattered into the
6x6-block BSR matrix, BCs applied.  `value` = element Jacobians per second
over the timed steps (inputs resident in HBM).

The same line reports, as separate fields:
  * the roofline of the element-Jacobian kernel (HIP events on the
    library's stream) and the BSR SpMV bandwidth on the assembled K;
  * `buckling_pipeline`: the 200 x 400 analysis of one design -- 5-step
    Newton on the nonlinear model, linear buckling about the final state
    (first 5 loads), KS (rho = 100) gradient of the loads with respect to
    all 80,000 thicknesses and 241,200 coordinates -- seconds per design;
  * `batched`: the 100 x 200 cylinder with 64 seeded thickness fields
    (seeds 2000..2063) solved in one batch -- linear static path, buckling
    (first 5 loads) and the KS thickness gradient -- seconds per design
    against the same analysis of one design alone;
  * `accuracy`: the first 5 loads of the 20 x 40 cylinder against the
    oracle (relative error), and the oracle's time for the same analysis;
  * `cpu_baseline`: the oracle's element loop on a bounded sample.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
                       [--skip-pipeline] [--skip-batched] [--skip-accuracy]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "a2d-shells_amd"))

import numpy as np  # noqa: E402

import shellgpu as sg  # noqa: E402
from shellgpu.problems import BoxBeam, Cylinder, seeded_imperfection, seeded_thickness  # noqa: E402

METRIC = "elem Jacobians/s, BSR GB/s; s/design (buckling+grad) vs plain GPU, λ err<1e-8"

# Floating-point operations of one element of the reference algorithm
# (TACSShellElement addJacobian / getMatType / addResidual), counted by
# compiling the oracle's operation-by-operation restatement with an
# operation-counting scalar (oracle/flopcount.cpp; +,-,*,/,sqrt = 1 flop).
FLOPS_PER_ELEMENT = {
    "stiffness_linear": 143931,
    "tangent_nonlinear": 155155,
    "geometric_fd": 310465,
    "residual_linear": 13651,
    "residual_nonlinear": 13831,
    "adjoint_thickness": 13781,
}
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBPS = 8000.0    # MI355X HBM3E
KS_RHO = 100.0


def progress(leg):
    """One line on stderr per finished leg (stdout carries only the JSON)."""
    print(f"[bench] {leg} done", file=sys.stderr, flush=True)


def spmv_bytes(nnzb, nnodes):
    """Algorithmic bytes of one 6x6-block SpMV: values (288 B/block), column
    indices (4 B/block), row pointers, x gathered once and y written once
    (48 B/node each)."""
    return 292 * nnzb + 4 * (nnodes + 1) + 96 * nnodes


def profiled_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3 summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration recorded in the same file); None if absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


# autotune each context (--no-tune keeps the default schedules: profiling
# runs, where the tuner's trial launches would fill the kernel tables)
TUNE = True


def make_context(cyl, thick, tune=None):
    """A context of the cylinder with its designs; the element-kernel
    schedules and the colouring are autotuned on it (outside every timed
    region), as a TACSAssembler is set up once per structure."""
    if tune is None:
        tune = TUNE
    ctx = sg.Context(0)
    ctx.set_mesh(cyl.conn, cyl.Xpts)
    ctx.set_material(70e9, 0.3, 2700.0, 0.0, 10.0)
    ctx.set_bcs(*cyl.bcs())
    ctx.set_designs(thick)
    ctx.set_ordering_seed(cyl.ring_nodes(0))
    if tune:
        ctx.tune(coloring=True)
    return ctx


def ks_of_loads(lam):
    """KS aggregate (rho = 100) of -lambda_i / lambda_1 for each design (a
    smooth minimum of the loads, normalised by the first, the normalisation
    differentiated too): value and d KS / d lambda."""
    ks, coef = [], []
    for row in lam:
        v, c = sg.ks_aggregate(-row / row[0], KS_RHO)
        d = -c / row[0]
        d[0] += float(c @ row) / row[0] ** 2
        ks.append(v)
        coef.append(d)
    return np.array(ks), np.array(coef)


def timed(ctx, fn):
    ctx.synchronize()
    t0 = time.perf_counter()
    out = fn()
    ctx.synchronize()
    return out, time.perf_counter() - t0


def buckling_pipeline(newton=True, coords=True):
    """200 x 400: Newton (5 load steps, nonlinear model) + buckling about the
    final state + KS gradient (thickness and coordinates); one design."""
    cyl = Cylinder(200, 400, R=1.0, L=2.0, imperfection=seeded_imperfection(1235))
    thick = seeded_thickness(cyl.nelems, 1234)
    ctx = make_context(cyl, thick)
    f = cyl.edge_load(1.0)
    t = {}
    u = None
    # the analysis objects are made once per structure, outside the timed
    # region; Newton's factor (its tangent at the zero state = K) is shared
    # with the buckling analysis, whose automatic shift needs K^-1
    F = ctx.factor() if newton else None
    B = ctx.buckling(sigma=0.0, nev=5, max_lanczos=150, tol=1e-10, model=sg.NONLINEAR if newton else sg.LINEAR)
    if newton:
        (u, its, rel), t["newton_s"] = timed(
            ctx, lambda: ctx.newton_solve(f, nsteps=5, max_iters=20, rtol=1e-10, factor=F))
        B.set_stiffness_factor(F)
    _, t["buckling_s"] = timed(ctx, lambda: B.solve(force=None if newton else f, u0=u))
    lam, err = B.eigenvalues()
    steps = B.info()["steps"]
    ks, coef = ks_of_loads(lam)
    (gt, gX), t["gradient_s"] = timed(ctx, lambda: B.eigen_sens(coef, thickness=True, coords=coords))
    B.destroy()
    if F is not None:
        F.destroy()
    ctx.close()
    total = sum(t.values())
    out = {"workload": "cylinder 200x400, 1 design: 5-step Newton (nonlinear model), buckling about the final "
                       "state (5 loads), KS(rho=100) gradient w.r.t. 80,000 thicknesses"
                       + (" and 241,200 coordinates" if coords else ""),
           "s_per_design": total, **t, "lanczos_steps": steps, "lambda": lam[0].tolist(),
           "max_ritz_residual": float(err.max()), "ks": float(ks[0]),
           "grad_t_norm": float(np.linalg.norm(gt)), "grad_X_norm": float(np.linalg.norm(gX)) if coords else None}
    if newton:
        out["newton_iterations"] = its
        out["newton_relres"] = float(rel[0])
    return out


def frequency_analysis():
    """200 x 400 (seeded thickness): mass and stiffness assembly, factor of K,
    Lanczos for the first 5 natural frequencies, their thickness gradient."""
    cyl = Cylinder(200, 400, R=1.0, L=2.0, imperfection=seeded_imperfection(1235))
    thick = seeded_thickness(cyl.nelems, 1234)
    ctx = make_context(cyl, thick)
    Fr = ctx.frequency(sigma=0.0, nev=5, max_lanczos=150, tol=1e-10)
    _, solve_s = timed(ctx, Fr.solve)
    lam, err = Fr.eigenvalues()
    steps = Fr.info()["steps"]
    gt, grad_s = timed(ctx, lambda: Fr.eigen_sens(np.ones(5)))
    Fr.destroy()
    ctx.close()
    return {"workload": "cylinder 200x400, 1 design: K and M, first 5 natural frequencies (shift-invert "
                        "Lanczos in the M inner product), thickness gradient of their sum",
            "solve_s": solve_s, "gradient_s": grad_s, "lanczos_steps": steps,
            "omega2": lam[0].tolist(), "freq_hz": (np.sqrt(lam[0]) / (2 * np.pi)).tolist(),
            "max_ritz_residual": float(err.max()), "grad_t_norm": float(np.linalg.norm(gt))}


def ks_of_positive_loads(lam, npos=5):
    """KS (rho = 100) of -lambda_i / lambda_1 over the first npos positive
    loads of each load case (a bent beam also has negative loads, buckling
    under the reversed load), as coefficients on the full eigenvalue list."""
    ks, coef = [], np.zeros_like(lam)
    for d, row in enumerate(lam):
        idx = np.where(row > 0)[0][:npos]
        v, c = ks_of_loads(row[idx][None, :])
        ks.append(v[0])
        coef[d, idx] = c[0]
    return np.array(ks), coef


def box_beam(ncases=16, chunk=16):
    """Stiffened box beam (128,000 quads, 746,352 dof): 16 load cases (tip
    force / bottom-skin pressure mixes) batched `chunk` at a time (one
    batch: the load cases share one factor of K): static path, buckling
    (first loads), KS thickness gradient per load case."""
    b = BoxBeam()
    t = b.panel_thickness(4000)
    t0 = time.perf_counter()
    ctx = sg.Context(0)
    ctx.set_mesh(b.conn, b.Xpts)
    ctx.set_material(70e9, 0.3)
    ctx.set_bcs(*b.bcs())
    ctx.set_designs(np.tile(t, (chunk, 1)))
    ctx.set_ordering_seed(b.root_nodes())
    if TUNE:
        ctx.tune(coloring=True)
    ctx.synchronize()
    setup_s = time.perf_counter() - t0
    lam_all, err_all, gpan, steps = [], [], [], []

    # allocated once (factors, Lanczos basis), reused by every batch
    B = ctx.buckling(sigma=0.0, nev=8, max_lanczos=150, tol=1e-10)

    def run(forces):
        B.solve(force=forces)
        lam, err = B.eigenvalues()
        _, coef = ks_of_positive_loads(lam)
        gt = B.eigen_sens(coef)[0]
        st = B.info()["steps"]
        return lam, err, gt, st

    total = 0.0
    for c0 in range(0, ncases, chunk):
        forces = np.stack([b.load_case(c, ncases) for c in range(c0, c0 + chunk)])
        (lam, err, gt, st), sec = timed(ctx, lambda: run(forces))
        total += sec
        lam_all.append(lam)
        err_all.append(err)
        gpan.append(np.stack([np.bincount(b.panel_of, weights=g, minlength=b.npanels) for g in gt]))
        steps.append(st)
    B.destroy()
    ctx.close()
    lam = np.concatenate(lam_all)
    first = np.array([row[row > 0][0] for row in lam])
    return {"workload": f"stiffened box beam (skins 2 x 400x80, spars 2 x 400x16, 40 ribs 80x16: "
                        f"{b.nelems} quads, {6 * b.nnodes} dof), {ncases} load cases batched {chunk} at a "
                        "time: static path, buckling (8 loads by magnitude), KS(rho=100) gradient of the "
                        "positive loads w.r.t. element (and 200 panel) thicknesses",
            "s_total": total, "s_per_load_case": total / ncases, "setup_s": setup_s,
            "lanczos_steps": steps, "first_positive_load": first.tolist(),
            "max_ritz_residual": float(np.concatenate(err_all).max()),
            "panel_grad_norm": float(np.linalg.norm(np.concatenate(gpan)))}


def batched_designs(nb=64):
    """100 x 200, nb thickness fields (seeds 2000..): linear path, buckling
    (5 loads), KS thickness gradient -- batched, against one design alone."""
    cyl = Cylinder(100, 200, R=1.0, L=2.0)
    thick = np.concatenate([seeded_thickness(cyl.nelems, 2000 + d) for d in range(nb)])
    f = cyl.edge_load(1.0)

    def run(th):
        # the buckling object (factors, Lanczos basis) is allocated once per
        # structure, as a TACSLinearBuckling is; the timed analysis reuses it
        ctx = make_context(cyl, th)
        B = ctx.buckling(sigma=0.0, nev=5, max_lanczos=100, tol=1e-10)

        def analysis():
            B.solve(force=f)
            lam = B.eigenvalues()[0]
            _, coef = ks_of_loads(lam)
            gt = B.eigen_sens(coef, thickness=True)[0]
            return lam, gt
        (lam, gt), sec = timed(ctx, analysis)
        B.destroy()
        ctx.close()
        return lam, gt, sec

    lam1, gt1, s1 = run(thick[:1])
    lam1b, gt1b, s1b = run(thick[:1])  # warm
    lamb, gtb, sb = run(thick)
    single = min(s1, s1b)
    return {"workload": f"cylinder 100x200, {nb} thickness designs batched: static path, buckling (5 loads), "
                        "KS(rho=100) thickness gradient",
            "designs": nb, "batched_s": sb, "batched_s_per_design": sb / nb, "single_s_per_design": single,
            "speedup_per_design": single / (sb / nb),
            "batched_vs_single_lambda_max_rel_diff": float(np.abs(lamb[0] - lam1[0]).max() / abs(lam1[0]).max()),
            "batched_vs_single_grad_max_rel_diff": float(np.abs(gtb[0] - gt1[0]).max() / np.abs(gt1[0]).max())}


def spmv_hbm(nb=64, reps=20):
    """6x6-block SpMV over 64 matrices of the 100 x 200 cylinder (one batched
    launch, 64 x 52 MB = 3.3 GB of values: far beyond the 256 MB MALL, so
    the matrix streams from HBM)."""
    cyl = Cylinder(100, 200, R=1.0, L=2.0)
    thick = np.concatenate([seeded_thickness(cyl.nelems, 2000 + d) for d in range(nb)])
    ctx = make_context(cyl, thick)
    K = ctx.matrix()
    ctx.assemble_mattype(sg.STIFFNESS, K)
    x = ctx.vector(np.random.default_rng(8).standard_normal(nb * 6 * cyl.nnodes))
    y = ctx.vector()
    for _ in range(3):
        K.mult(x, y)
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(reps):
        K.mult(x, y)
    ms, n = ctx.timer_stop(), reps
    per = spmv_bytes(ctx.nnzb, ctx.nnodes) * nb
    out = {"workload": f"cylinder 100x200, {nb} stiffness matrices (one batched launch)",
           "bytes_per_launch": per, "avg_launch_ms": ms / n, "GBps": per / (ms / n * 1e-3) / 1e9,
           "frac_of_hbm_peak": per / (ms / n * 1e-3) / 1e9 / PEAK_HBM_GBPS}
    K.destroy()
    ctx.close()
    return out


def accuracy_vs_oracle():
    """First 5 loads of the 20 x 40 cylinder (t = 0.002) against the oracle;
    both run the same shift-invert Lanczos problem."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    cyl = Cylinder(20, 40)
    thick = np.full((1, cyl.nelems), 0.002)
    sigma = 0.9 * cyl.classical_load()
    f = cyl.edge_load(1.0)
    ctx = make_context(cyl, thick)
    out, gpu_s = timed(ctx, lambda: ctx.linear_buckling(sigma, f, nev=8, max_lanczos=120, tol=1e-12))
    ctx.close()
    t0 = time.perf_counter()
    lam = O.linear_buckling(cyl.Xpts, cyl.conn, thick[0], cyl.bcs(), f, sigma, nev=8, max_lanczos=120)[0]
    cpu_s = time.perf_counter() - t0
    g = out["lambda"][0]
    return {"workload": "cylinder 20x40, t = 0.002: static path + buckling, first 5 loads",
            "lambda_max_rel_err": float(np.abs(g[:5] - lam[:5]).max() / np.abs(lam[:5]).max()),
            "gpu_s": gpu_s, "oracle_cpu_s": cpu_s, "oracle_cores": 1}


def host_cpu():
    """Host CPU model and logical CPU count (from /proc/cpuinfo)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def cpu_workers():
    """Host worker processes for the CPU baselines: every logical CPU, at
    most 16 (the CPU share of one GPU on the measurement hosts)."""
    return max(1, min(16, os.cpu_count() or 1))


def _oracle_k_worker(args):
    """One worker of the oracle's element K loop (BSR scatter included)."""
    Xpts, conn, thick, nnodes, seconds = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    m = O.material()
    rowp, cols = O.node_csr(nnodes, conn)
    n_done, t_used = 0, 0.0
    while t_used < seconds:
        t0 = time.perf_counter()
        O.assemble_mat(O.KIND_K, 0, m, Xpts, conn, thick, rowp, cols)
        t_used += time.perf_counter() - t0
        n_done += conn.shape[0]
    return n_done, t_used


def cpu_baseline(cyl, thick, seconds):
    """The reference's own element code on the host's cores, when its build
    is present (oracle/_ref/ref_timer: TACSQuad4Shell::getMatType(K) compiled
    from the reference's sources), one process per worker; the oracle's
    element loop with its BSR scatter (a scalar C restatement of the
    reference) is reported beside it, and is the baseline where the
    reference build is absent.  Each worker repeats K of the same 2,000
    elements for ~`seconds`; the rate is the sum over workers."""
    import multiprocessing as mp
    import subprocess
    import tempfile

    nw = cpu_workers()
    sample = 2000
    conn = np.ascontiguousarray(cyl.conn[:sample])
    with mp.get_context("fork").Pool(nw) as pool:
        res = pool.map(_oracle_k_worker, [(cyl.Xpts, conn, thick[0, :sample], cyl.nnodes, seconds)] * nw)
    port = {
        "value": sum(d / t for d, t in res),
        "unit": "elem Jacobians/s",
        "cores": nw,
        "kind": "port",
        "sample": f"K of the first {sample} elements of the 200x400 cylinder assembled into BSR by the oracle, "
                  f"repeated for ~{seconds:.0f} s in each of {nw} processes ({sum(d for d, _ in res)} element "
                  "Jacobians)",
        "host": host_cpu(),
    }
    timer = os.path.join(ROOT, "oracle", "_ref", "ref_timer")
    if not os.path.exists(timer):
        return port
    rec = np.concatenate([cyl.Xpts[conn].reshape(sample, 12), thick[0, :sample, None]], axis=1)
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as fh:
        rec.astype(np.float64).tofile(fh)
        path = fh.name
    try:
        procs = [subprocess.Popen([timer, path, str(seconds)], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                                  text=True) for _ in range(nw)]
        outs = [p.communicate(timeout=seconds + 120)[0].split() for p in procs]
        rates = [int(o[0]) / float(o[1]) for o in outs]
        done = sum(int(o[0]) for o in outs)
    except (subprocess.SubprocessError, ValueError, OSError, IndexError):
        return port
    finally:
        os.unlink(path)
    return {
        "value": sum(rates),
        "unit": "elem Jacobians/s",
        "cores": nw,
        "kind": "reference",
        "sample": f"TACSQuad4Shell::getMatType(TACS_STIFFNESS_MATRIX) (the reference's element code, compiled "
                  f"from its sources) on the first {sample} elements of the 200x400 cylinder, repeated for "
                  f"~{seconds:.0f} s in each of {nw} processes ({done} element Jacobians; one process per core); "
                  "its assembler (needs METIS) is not built, so the scatter into BSR is not in this figure",
        "per_core": min(rates),
        "host": host_cpu(),
        "port": {"value": port["value"], "cores": nw, "sample": port["sample"]},
    }


def _oracle_analysis_worker(_):
    """One oracle analysis of the 20 x 40 cylinder (static path + buckling)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    cyl = Cylinder(20, 40)
    thick = np.full(cyl.nelems, 0.002)
    t0 = time.perf_counter()
    O.linear_buckling(cyl.Xpts, cyl.conn, thick, cyl.bcs(), cyl.edge_load(1.0), 0.9 * cyl.classical_load(), nev=8,
                      max_lanczos=120)
    return time.perf_counter() - t0


def cpu_analysis_throughput():
    """The oracle's whole 20 x 40 analysis on every worker at once: analyses
    per second of the host (to set beside the GPU's single-analysis time)."""
    import multiprocessing as mp
    nw = cpu_workers()
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(nw) as pool:
        per = pool.map(_oracle_analysis_worker, range(nw))
    wall = time.perf_counter() - t0
    return {"workers": nw, "analyses_per_s": nw / wall, "s_per_analysis_one_core": float(np.median(per)),
            "host": host_cpu()}


def stiffness_kernel_label(schedule):
    """The stiffness kernel the timed assembly ran: its tuned schedule."""
    elem = schedule.get("elem", {}).get("stiffness")
    if elem is None:
        return "elem_kernel<0> (stiffness), default schedule 0: " + "two elements per workgroup"
    v = elem["schedule"]
    return f"elem_kernel<0> (stiffness), tuned schedule {v}: {schedule['schedules'][v]}" + (
        "" if v & 4 else ", two elements per workgroup")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--skip-pipeline", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-accuracy", action="store_true")
    ap.add_argument("--skip-frequency", action="store_true")
    ap.add_argument("--skip-boxbeam", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default schedules (profiling runs)")
    ap.add_argument("--skip-spmv-hbm", action="store_true",
                    help="no HBM-resident SpMV leg (its 64-design assembly runs the same element kernel)")
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--mode", choices=["assembly", "calibrate"], default="assembly",
                    help="calibrate: only run a streaming kernel of known traffic (Y += a X over "
                         "3 x 36 x nnzb doubles of the workload's matrices) for counter calibration")
    args = ap.parse_args()
    global TUNE
    TUNE = not args.no_tune
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only; a decomposition across GPUs is not a "
                                   "measured path of this library", "n_gpus": args.gpus}))
        sys.exit(2)

    cyl = Cylinder(200, 400, R=1.0, L=2.0, imperfection=seeded_imperfection(1235))
    thick = seeded_thickness(cyl.nelems, 1234)
    # host baselines first: their worker processes fork a process that has
    # not touched the GPU
    cpu = None
    cpu_analysis = None
    if args.mode == "assembly" and not args.no_cpu_baseline:
        cpu = cpu_baseline(cyl, thick, args.cpu_seconds)
        progress("cpu_baseline")
        if not args.skip_accuracy:
            cpu_analysis = cpu_analysis_throughput()
            progress("cpu_analysis")
    ctx = make_context(cyl, thick)
    K = ctx.matrix()

    if args.mode == "calibrate":
        K2 = ctx.matrix()
        n = 36 * ctx.nnzb
        for _ in range(3):
            K2.axpy(0.5, K)  # reads 2n doubles, writes n doubles
        ctx.synchronize()
        print(json.dumps({"calibrate": "axpy", "doubles": n, "read_bytes": 16 * n, "write_bytes": 8 * n,
                          "launches": 3}), flush=True)
        return

    for _ in range(args.warmup):
        ctx.assemble_mattype(sg.STIFFNESS, K)
    ctx.synchronize()

    ctx.set_timing(True)
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.assemble_mattype(sg.STIFFNESS, K)
    wall_ms = ctx.timer_stop()
    elem_ms, elem_launches, _, _ = ctx.get_timing()
    ctx.set_timing(False)

    n_jac = cyl.nelems * args.steps
    value = n_jac / (wall_ms * 1e-3)
    flops = FLOPS_PER_ELEMENT["stiffness_linear"] * n_jac
    achieved = flops / (elem_ms * 1e-3) / 1e12

    # SpMV on the assembled K
    x = ctx.vector(np.random.default_rng(7).standard_normal(6 * cyl.nnodes))
    y = ctx.vector()
    for _ in range(3):
        K.mult(x, y)
    # one pair of stream events around the back-to-back products (an event
    # pair around each ~35 us product adds a few us to it)
    nspmv = 50
    ctx.synchronize()
    ctx.timer_start()
    for _ in range(nspmv):
        K.mult(x, y)
    spmv_ms = ctx.timer_stop()
    spmv_gbps = spmv_bytes(ctx.nnzb, ctx.nnodes) * nspmv / (spmv_ms * 1e-3) / 1e9
    nnzb, nnodes, ncolors = ctx.nnzb, ctx.nnodes, ctx.ncolors
    device = ctx.device_info()
    schedule = ctx.schedule()
    K.destroy()
    x.free()
    y.free()
    ctx.close()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "elem Jacobians/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded cylinder mesh, thickness and imperfection)",
        "config": {"workload": "cylinder 200x400 (80000 quads, 482400 dof): K assembly", "nelems": cyl.nelems,
                   "nnodes": nnodes, "nnzb": nnzb, "colors": ncolors, "parallelism": "single GPU"},
        "device": device,
        "schedule": schedule,
        "bsr_spmv_GBps": spmv_gbps,
        "bsr_spmv_note": "200x400 K (207.7 MB) fits the 256 MB MALL: consecutive products partly hit it; "
                         "bsr_spmv_hbm is the HBM-resident figure",
        "roofline": {
            # vector FP64 FMAs (the element tile is far below an efficient
            # FP64 MFMA shape); the peak is MI355X's FP64 rate
            "bound": "fp64-valu",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": profiled_traffic("elem_stiffness_cylinder200x400.md"),
            "kernel": stiffness_kernel_label(schedule),
            "avg_launch_ms": elem_ms / max(elem_launches, 1),
            "flops_per_element": FLOPS_PER_ELEMENT["stiffness_linear"],
        },
    }
    legs = [("bsr_spmv_hbm", not args.skip_spmv_hbm, spmv_hbm),
            ("buckling_pipeline", not args.skip_pipeline, buckling_pipeline),
            ("frequency", not args.skip_frequency, frequency_analysis),
            ("box_beam", not args.skip_boxbeam, box_beam),
            ("batched", not args.skip_batched, lambda: batched_designs(args.batch)),
            ("accuracy", not args.skip_accuracy, accuracy_vs_oracle)]
    progress("assembly")
    for name, run, fn in legs:
        if run:
            out[name] = fn()
            progress(name)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    if cpu_analysis is not None and "accuracy" in out:
        acc = out["accuracy"]
        acc["oracle_all_cores"] = cpu_analysis
        acc["gpu_analyses_per_s"] = 1.0 / acc["gpu_s"]
        acc["gpu_vs_all_cores"] = acc["gpu_analyses_per_s"] / cpu_analysis["analyses_per_s"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
