"""Benchmark of the MAS-PCG Newton-step linear solve on one MI355X.

Headline workload: 64 independent FEM + contact systems solved together
(cubes whose vertex counts are drawn uniformly from 5k to 80k, seed 0:
2,762,169 vertices, 8.29 M DOF; stable Neo-Hookean Hessians with lognormal E,
5% of each cube's vertices in point-triangle contact pairs, 32-vertex Morton
subdomains, reference level rule), tolerance 1e-4 on the FP64 true relative
residual of every system.  Synthetic data, generated in-process.

One step = one Newton step's linear solves with fixed structure: the new
matrix values (already in HBM, caller layout) are taken in, every system's
MAS preconditioner is rebuilt from them (FP64 Galerkin + factorization,
inverse stored in the chosen precision) and the batched PCG runs until every
system meets the tolerance.  `value` is the plain FP32 step time; the
mixed-precision (FP16 / BF16 inverse storage) times and their speed-ups are
reported beside it, as are the 40^3 + contacts, hybrid ABD + FEM and 60^3
adjoint workloads.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps of the headline (64-system) workload")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--precisions", default="fp32,fp16,bf16",
                    help="comma list; fp32 is always measured (it is the headline value)")
    ap.add_argument("--kernel-reps", type=int, default=20, help="launches per kernel timing window (batch)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines (and the oracle solution errors)")
    ap.add_argument("--cpu-runs", type=int, default=3, help="timed repetitions of each CPU baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline workload (no 40^3, hybrid, adjoint or structure-change legs)")
    ap.add_argument("--batch-only", action="store_true",
                    help="the headline workload's GPU figures only, no CPU leg (profiling aid)")
    ap.add_argument("--fem40-only", action="store_true",
                    help="only the 40^3 + contacts GPU figures (profiling aid)")
    ap.add_argument("--hybrid-only", action="store_true", help="only the hybrid GPU figures (profiling aid)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint-differentiated solve and its figures (profiling aid)")
    ap.add_argument("--n", type=int, default=40, help="hex cells per edge of the single-system workload")
    ap.add_argument("--contacts", type=int, default=20000)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--fem40-steps", type=int, default=20)
    ap.add_argument("--batch-tiling", type=int, default=0, choices=(0, 2, 4, 8),
                    help="SpMV entries per thread on the batch; 0 times every tiling and keeps the fastest")
    ap.add_argument("--spmv-tiling", type=int, default=0, choices=(0, 2, 4, 8),
                    help="SpMV entries per thread on the 40^3 workload; 0 autotunes")
    ap.add_argument("--adjoint-n", type=int, default=60, help="hex cells per edge of the adjoint workload")
    ap.add_argument("--side-steps", type=int, default=5, help="timed steps of the hybrid and adjoint legs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (HBM counter calibration) and exit")
    return ap.parse_args()


def algorithmic_bytes(info, n_vertices, prec_bytes):
    """Bytes one launch of each iteration kernel must move at minimum.

    spmv: nb*(36+4) values+indices of the upper blocks, (n+1)*4 row pointers,
          12 B/row x and 12 B/row y, plus the fused p = z + beta p update
          (z and old p read instead of x: +12 B/row, p written: +12 B/row).
    mas_local: every bank inverse (528 blocks x 9 x s) and the bank's R read
          and Z written (32 slots x 12 B each way).
    update: x, r, p, Ap read, x and r written (6 x 12 B/row); level-1
          restriction output is folded into the bank term.
    precond_finish: r and z read, z written (3 x 12 B/row)."""
    n = n_vertices
    nb = info["nb_total"]
    banks = info["n_banks"]
    bank = info["bank"]
    nblk = bank * (bank + 1) // 2
    return {
        "spmv": nb * 40 + (n + 1) * 4 + n * 24 + n * 24,
        "mas_local": banks * nblk * 9 * prec_bytes + banks * bank * 12 * 2,
        "update": n * 12 * 6,
        "precond_finish": n * 12 * 3,
    }


KERNEL_PREFIX = {"spmv": "k_spmv_fused", "mas_local": "k_mas_local<32, float,", "update": "k_update",
                 "precond_finish": "k_precond_finish"}


def measured_traffic(kernel, workload_is_default, spmv_entries, summary="fem40c_hbm_traffic.json", waves=None,
                     spmv_nt=False):
    """HBM bytes per launch of `kernel` (the SpMV at `spmv_entries` per
    thread with nontemporal matrix loads or not, the FP32 MAS local kernel at
    `waves` waves per workgroup) from the
    committed counter summary of this workload (FETCH_SIZE + WRITE_SIZE runs,
    copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", summary)
    if not workload_is_default or not os.path.exists(path):
        return None
    data = json.load(open(path))["kernels"]
    want = (f"k_spmv_fused<{spmv_entries}, {'true' if spmv_nt else 'false'}," if kernel == "spmv"
            else KERNEL_PREFIX[kernel])
    if kernel == "mas_local" and waves:
        from agipc_learning_amd.prof_summary import mas_local_variant
        want = f"{KERNEL_PREFIX[kernel]} {mas_local_variant(waves)}>"
    for name, v in data.items():
        if name.startswith(want):
            return int(v["bytes"])
    return None


def rhs_variants(b, count=4, seed=7):
    """The right-hand side and count - 1 perturbed copies (each entry scaled by
    1 + 0.2 N(0, 1), seeded): consecutive timed steps solve different systems
    b, so a solve cannot reuse the previous one's iteration count exactly
    (the solver launches chunks ahead from it); the last timed step solves
    the original b."""
    import torch
    gen = torch.Generator(device=b.device).manual_seed(seed)
    out = [b]
    for _ in range(count - 1):
        out.append(b * (1.0 + 0.2 * torch.randn(b.shape, generator=gen, device=b.device, dtype=b.dtype)))
    return out


def kernel_times(s, reps):
    """Per-kernel execution time (us) of the four iteration kernels: events
    stamped by each launch's own dispatch, on the library's stream.  Each
    kernel gets five windows of `reps` launches, taken round-robin over the
    kernels with 10 ms between rounds so the windows of one kernel lie apart
    in time, and keeps its lowest window mean (a slow episode of some ms on
    the box then spoils a window or two, not the figure)."""
    names = ["mas_local", "precond_finish", "spmv", "update"]
    best = {n: float("inf") for n in names}
    for rnd in range(5):
        if rnd:
            time.sleep(0.01)
        for k, n in enumerate(names):
            best[n] = min(best[n], s.time_kernel(k, reps) * 1e3)
    return best


def copy_calibration():
    import torch

    n = 256 << 20
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 5}))


def stream_ceilings():
    """Practical memory rates of this card from the library's own streaming
    kernels (masx_stream_rate): a read-only stream and a copy over 4 GiB (far
    beyond the 256 MB Infinity Cache), fastest of 10 launches, GB/s."""
    from agipc_learning_amd import native

    n = 4 << 30
    return {"read_GBs": round(native.stream_rate(n, "read", 10), 1),
            "copy_GBs": round(native.stream_rate(n, "copy", 10), 1), "bytes": n}


def host_cpu():
    """What the CPU legs run on: logical CPUs of the host, the CPUs this
    process may use, and the model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def oracle_module():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    O.build()
    return O


def oracle_setup(O, w, levels=0):
    """The oracle's MAS with its hierarchy (outside the timed region, as the
    GPU step reuses its structure analysis too)."""
    import numpy as np

    nv = w["n_vertices"]
    p2r, r2p = O.morton_partition(w["positions"])
    start, num, lst = O.mesh_neighbors(nv, w["tets"])
    mas = O.Mas(nv, len(p2r), start, num, lst, p2r, r2p, levels=levels)
    mas.reorder(w["pairs"])
    fem = (w["rows"] >= w["fem_offset"]) & (w["cols"] >= w["fem_offset"])
    idx = None if fem.all() else np.nonzero(fem)[0].astype(np.uint32)
    return mas, idx


def oracle_step(O, w, mas, idx, tol):
    """One step of the oracle (FP64 scalar C restatement of the reference
    algorithm): MAS assembly + Cholesky-built inverse + PCG to the tolerance
    (ABD 12x12 block-Jacobi on the bodies when there are any)."""
    mas.assemble(w["vals"], w["rows"], w["cols"], indices=idx, offset=w["fem_offset"])
    mas.invert(1)
    abd = O.abd_block_inverse(w["n_bodies"], w["vals"], w["rows"], w["cols"]) if w["n_bodies"] else None
    return O.pcg(w["block_rows"], w["vals"], w["rows"], w["cols"], w["rhs"], n_abd=w["n_bodies"], abd_inv=abd,
                 mas=mas, fem_offset=w["fem_offset"], tol=tol, rule=1)


def stats_of(times_ms):
    t = sorted(times_ms)
    return {"median": round(t[len(t) // 2], 2), "min": round(t[0], 2), "max": round(t[-1], 2), "runs": len(t)}


def cpu_batch(O, ws, tol, runs):
    """The oracle on the whole 64-system batch on every thread it has, `runs`
    times (each run = one step of all 64 systems: assembly + inverse + PCG),
    and on one thread on a sample of the systems, scaled to the batch by
    vertex count.  Also returns the oracle's solutions."""
    threads = O.get_threads()
    per_run = [0.0] * runs
    sols, iters = [], []
    single_sample = list(range(0, len(ws), 16))  # 4 of the 64 systems
    single = [0.0] * runs
    for i, w in enumerate(ws):
        mas, idx = oracle_setup(O, w)
        for r in range(runs):
            t0 = time.perf_counter()
            x, so = oracle_step(O, w, mas, idx, tol)
            per_run[r] += (time.perf_counter() - t0) * 1e3
        sols.append(x)
        iters.append(so["iterations"])
        if i in single_sample:
            O.set_threads(1)
            for r in range(runs):
                t0 = time.perf_counter()
                oracle_step(O, w, mas, idx, tol)
                single[r] += (time.perf_counter() - t0) * 1e3
            O.set_threads(threads)
        del mas
        if i % 16 == 15:
            log(f"  cpu batch: {i + 1} systems")
    nv = sum(w["n_vertices"] for w in ws)
    nv_s = sum(ws[i]["n_vertices"] for i in single_sample)
    scale = nv / nv_s
    all_core = stats_of(per_run)
    one = stats_of([t * scale for t in single])
    return {"value": all_core["median"], "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"the whole 64-system step ({nv} vertices): MAS assembly + Cholesky-built inverse + PCG to "
                      f"{tol:g} per system, FP64 oracle (scalar C restatement of the reference algorithm, OpenMP "
                      f"loops over banks and rows on {threads} threads), {runs} runs; hierarchy built before the "
                      "timed region",
            "all_cores": dict(all_core, threads=threads),
            "single_core": dict(one, threads=1, sample=f"systems {single_sample} ({nv_s} vertices) timed on one "
                                f"thread, scaled to the batch by vertex count (x{scale:.2f})"),
            "host": host_cpu()}, sols, iters


def cpu_single(O, w, levels, tol, runs):
    """The oracle on one step of a single-system workload, every thread and
    one thread, `runs` times each; returns the oracle's solution too."""
    mas, idx = oracle_setup(O, w, levels)
    threads = O.get_threads()
    times, x, so = [], None, None
    for _ in range(runs):
        t0 = time.perf_counter()
        x, so = oracle_step(O, w, mas, idx, tol)
        times.append((time.perf_counter() - t0) * 1e3)
    O.set_threads(1)
    one = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle_step(O, w, mas, idx, tol)
        one.append((time.perf_counter() - t0) * 1e3)
    O.set_threads(threads)
    all_core = stats_of(times)
    return {"value": all_core["median"], "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"one full step (MAS assembly + Cholesky-built inverse + PCG, {so['iterations']} iterations, "
                      f"true rel res {so['true_rel_res']:.2e}), FP64 oracle, {runs} runs",
            "all_cores": dict(all_core, threads=threads), "single_core": dict(stats_of(one), threads=1),
            "oracle_iterations": so["iterations"]}, x


def rel_err(a, b):
    import numpy as np

    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def batch_leg(args):
    """The headline: 64 independent cube systems (vertex counts U[5k, 80k],
    seed 0; 5% of vertices in contact pairs) solved together in one batched
    launch sequence, plain FP32 vs FP16 / BF16 inverse storage.  One step =
    values in + build + solve of all 64 systems."""
    import numpy as np
    import torch
    from agipc_learning_amd import session
    from agipc_learning_amd import workloads as W

    log("step: batch64 generate")
    ws = W.batch64(seed=0)
    nv = sum(w["n_vertices"] for w in ws)
    log(f"step: batch64 analyze + setup ({nv} vertices, {sum(w['nb'] for w in ws)} upper blocks)")
    t0 = time.perf_counter()
    s = session.solver_for(ws)
    setup_ms = (time.perf_counter() - t0) * 1e3
    vals = [torch.from_numpy(np.ascontiguousarray(w["vals"])).cuda() for w in ws]
    b = torch.from_numpy(session.rhs_of(ws)).cuda()
    x = torch.zeros_like(b)
    info = s.info()
    s.build("fp32")
    if args.batch_tiling:
        s.set_spmv_tiling(args.batch_tiling)
        tiling = {"entries_per_thread": args.batch_tiling, "ms": s.time_kernel(2, 20),
                  "nontemporal": s.spmv_nontemporal()}
    else:
        log("step: batch64 autotune SpMV tiling")
        tiling = s.autotune(20)
    log(f"  tiling {tiling}")

    bs = rhs_variants(b)
    torch.cuda.synchronize()
    # the step runs on one stream, as a Newton loop would call it: values in
    # (all 64 systems in one call), build and solve are ordered on the stream;
    # the solve returns after its statistics are read
    work = torch.cuda.Stream()
    ws_ = work.cuda_stream

    all_vals = dict(enumerate(vals))

    def step(prec, k=0):
        s.set_values_many(all_vals, stream=ws_)
        s.build(prec, stream=ws_)
        return s.solve(bs[k % len(bs)], x, rule=1, tol=args.tol, stream=ws_)

    precs = ["fp32"] + [p for p in args.precisions.split(",") if p and p != "fp32"]
    out = {"systems": len(ws), "vertices": nv, "dof": 3 * nv, "upper_blocks": int(info["nb_total"]),
           "banks": int(info["n_banks"]), "setup_ms": round(setup_ms, 1), "spmv_tiling": tiling,
           "rhs_variants": len(bs)}
    sols = {}
    tiling["mas_local"] = {}
    for prec in precs:
        # the MAS local variant for this precision (VALU waves per workgroup, or MFMA)
        s.build(prec)
        tiling["mas_local"][prec] = s.autotune_mas(20)
        log(f"step: batch64 {prec} ({args.warmup} warmup + {args.steps} steps), local variant "
            f"{tiling['mas_local'][prec]}")
        for i in range(args.warmup):
            step(prec, i + 1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(args.steps):
            st = step(prec, args.steps - 1 - i)  # the last step solves the original b
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / args.steps
        sols[prec] = x.cpu().numpy().copy()
        worst = max(q["rel_res_true"] for q in st)
        fails = sum(1 for q in st if q["failure"] != 0 or q["status"] != 1)
        if fails or worst > args.tol * (1 + 1e-6):
            raise SystemExit(f"batch64 {prec}: {fails} systems failed, worst true rel res {worst:.3e}")
        its = [q["iterations"] for q in st]
        # split of the step: build alone, solve alone
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            s.build(prec)
        torch.cuda.synchronize()
        build_ms = (time.perf_counter() - t) * 1e3 / 3
        t = time.perf_counter()
        for _ in range(3):
            s.solve(b, x, rule=1, tol=args.tol)
        torch.cuda.synchronize()
        pcg_ms = (time.perf_counter() - t) * 1e3 / 3
        kern = {k: round(v, 2) for k, v in kernel_times(s, args.kernel_reps).items()}
        s.build(prec)  # time_kernel leaves the iteration state frozen
        alg = algorithmic_bytes(info, nv, 4 if prec == "fp32" else 2)
        out[prec] = {"ms": round(ms, 3), "build_ms": round(build_ms, 3), "pcg_ms": round(pcg_ms, 3),
                     "max_iterations": max(its), "mean_iterations": round(sum(its) / len(its), 2),
                     "iterations": its, "worst_rel_res_true": worst,
                     "dof_iterations_per_s": round(sum(3 * w["n_vertices"] * q["iterations"]
                                                       for w, q in zip(ws, st)) / (pcg_ms * 1e-3), 1),
                     "kernels_us": kern,
                     "kernels_GBs": {k: round(alg[k] / (kern[k] * 1e-6) / 1e9, 1) for k in kern}}
        if prec == "fp32":
            dom = max(kern, key=lambda k: kern[k])
            ach = alg[dom] / (kern[dom] * 1e-6) / 1e9
            out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                               "traffic": measured_traffic(dom, True, tiling["entries_per_thread"],
                                                           "batch64_hbm_traffic.json",
                                                           tiling["mas_local"]["fp32"]["waves"],
                                                           tiling["nontemporal"]),
                               "algorithmic_bytes": alg[dom], "avg_us": kern[dom],
                               "per_kernel": {k: {"us": kern[k], "GB/s": out[prec]["kernels_GBs"][k],
                                                  "algorithmic_bytes": alg[k]} for k in kern}}
        log(f"  batch64 {prec}: {ms:.2f} ms/step (build {build_ms:.2f}, pcg {pcg_ms:.2f}), iterations max "
            f"{max(its)}, worst true rel res {worst:.2e}, kernels us {kern}, GB/s {out[prec]['kernels_GBs']}")
    for prec in precs[1:]:
        out[prec]["speedup_vs_fp32"] = round(out["fp32"]["ms"] / out[prec]["ms"], 3)
    if not args.no_extra:
        # a structure change of every system (new rows/cols and contact pairs,
        # here the same ones: the analysis reruns in full): MAS hierarchies,
        # block CSR, Galerkin maps, bank entries and SpMV layout on the device
        # from host arrays ("ms") and from device-resident arrays
        # ("device_arrays_ms", no upload of the pattern), and for comparison
        # the host analysis
        log("step: batch64 structure change")
        restr = {}
        dev_rc = [(torch.from_numpy(np.ascontiguousarray(w["rows"], np.int32)).cuda(),
                   torch.from_numpy(np.ascontiguousarray(w["cols"], np.int32)).cuda()) for w in ws]
        for key, on_device, from_device in (("ms", True, False), ("device_arrays_ms", True, True),
                                            ("host_structure_ms", False, False)):
            s.set_device_structure(on_device)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for i, w in enumerate(ws):
                if from_device:
                    s.set_structure_device(i, *dev_rc[i])
                else:
                    s.set_structure(i, w["rows"], w["cols"])
                s.set_contacts(i, w["pairs"])
            s.setup()
            torch.cuda.synchronize()
            restr[key] = round((time.perf_counter() - t) * 1e3, 1)
            if s.device_structure_used() != on_device:
                raise SystemExit("batch structure analysis ran on the wrong side")
        s.set_device_structure(True)
        del dev_rc
        out["structure_change"] = restr
        log(f"  structure change: {restr}")
    del s, vals, b, x
    torch.cuda.empty_cache()
    return out, ws, sols


def hybrid(args, w):
    """ABD + FEM system: 2,000 affine bodies (dense SPD 12x12 diagonal blocks),
    30,000 body-pair couplings, a 20^3 FEM block (9,261 vertices) coupled to
    200 bodies; ABD 12x12 block-Jacobi + MAS on the FEM part.  Plain FP32 vs
    FP16 / BF16 inverse storage.  Returns the figures and the solutions."""
    import numpy as np
    import torch
    from agipc_learning_amd import session

    log("step: hybrid setup")
    s = session.solver_for(w)
    vals = torch.from_numpy(np.ascontiguousarray(w["vals"])).cuda()
    b = torch.from_numpy(session.rhs_of(w)).cuda()
    x = torch.zeros_like(b)
    out = {"bodies": int(w["n_bodies"]), "fem_vertices": int(w["n_vertices"]), "dof": int(len(b)),
           "upper_blocks": int(s.info()["nb_total"])}
    xs = {}
    torch.cuda.synchronize()  # inputs in place before the step's own stream reads them
    work = torch.cuda.Stream()  # the step on one stream, as in the batch leg
    ws_ = work.cuda_stream
    for prec in ("fp32", "fp16", "bf16"):
        def step():
            s.set_values(0, vals, stream=ws_)
            s.build(prec, stream=ws_)
            return s.solve(b, x, rule=1, tol=args.tol, stream=ws_)[0]
        step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.side_steps):
            st = step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / args.side_steps
        xs[prec] = x.cpu().numpy().copy()
        out[prec] = {"ms": round(ms, 4), "iterations": st["iterations"], "rel_res_true": st["rel_res_true"],
                     "failure": st["failure"]}
        if prec != "fp32":
            out[prec]["speedup_vs_fp32"] = round(out["fp32"]["ms"] / ms, 3)
        log(f"  hybrid {prec}: {ms:.3f} ms/step, iterations {st['iterations']}, true rel res {st['rel_res_true']:.2e}")
    return out, xs


def adjoint_leg(args):
    """Adjoint-differentiated solve: 60^3 grid (226,981 vertices, 1,296,000
    tets), per-tet Lame parameters lognormal (sigma 0.5) around E = 1e5 Pa,
    nu = 0.4; L = 1/2 ||x - x*||^2, x* ~ N(0, 1).  One step = forward
    (Stable Neo-Hookean assembly on the GPU, MAS build, MAS-PCG) + backward
    (adjoint MAS-PCG with the same preconditioner, gradient kernel) through
    the torch.autograd.Function.  The production gradient (both solves at the
    tolerance) is compared with the FP64-refined one (both solves to 1e-9)."""
    import numpy as np
    import torch
    from agipc_learning_amd import adjoint
    from agipc_learning_amd import workloads as W

    n = args.adjoint_n
    log(f"step: adjoint generate {n}^3 grid")
    w = W.fem_grid(n, young_sigma=0.5, seed=0)
    log(f"step: adjoint setup ({w['n_vertices']} vertices, {len(w['tets'])} tets)")
    t0 = time.perf_counter()
    system = adjoint.ElasticSystem.from_workload(w, tol=args.tol)
    setup_ms = (time.perf_counter() - t0) * 1e3
    rng = np.random.default_rng(0)
    xstar = torch.from_numpy(rng.standard_normal(3 * w["n_vertices"])).cuda()
    b = torch.from_numpy(np.ascontiguousarray(w["rhs"])).cuda()
    mu0 = torch.from_numpy(np.ascontiguousarray(w["mu"])).cuda()
    lam0 = torch.from_numpy(np.ascontiguousarray(w["lam"])).cuda()

    def step():
        mu = mu0.clone().requires_grad_(True)
        lam = lam0.clone().requires_grad_(True)
        x = adjoint.solve(mu, lam, system, b)
        loss = 0.5 * ((x - xstar) ** 2).sum()
        loss.backward()
        return mu.grad, lam.grad, float(loss.detach())

    step()
    torch.cuda.synchronize()
    reps = max(2, args.side_steps)
    t = time.perf_counter()
    for _ in range(reps):
        gm, gl, loss = step()
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t) * 1e3 / reps
    fwd, adj = dict(system.forward_stats), dict(system.adjoint_stats)
    # parts: assembly, build, each solve, gradient kernel (device events on
    # the stream the library calls here run on)
    ev = lambda: torch.cuda.Event(enable_timing=True)

    def timed(fn, k=3):
        a, z = ev(), ev()
        fn()
        a.record()
        for _ in range(k):
            fn()
        z.record()
        torch.cuda.synchronize()
        return a.elapsed_time(z) / k

    mu_d, lam_d = mu0.detach(), lam0.detach()
    assemble_ms = timed(lambda: system.fem.assemble(system.dminv, system.vol, system.F, mu_d, lam_d,
                                                    base=system.base, out=system.vals))
    t = time.perf_counter()
    for _ in range(3):
        system.build(mu_d, lam_d)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t) * 1e3 / 3
    t = time.perf_counter()
    for _ in range(3):
        x = system.solve(b)
    torch.cuda.synchronize()
    solve_ms = (time.perf_counter() - t) * 1e3 / 3
    y = system.solve(x - xstar)
    grad_ms = timed(lambda: system.param_grad(mu_d, lam_d, x, y))
    # refined reference gradient (both solves to 1e-9 by FP64 refinement; the
    # faster of two steps)
    system.tol = 1e-9
    refined_ms = 1e30
    for _ in range(2):
        torch.cuda.synchronize()
        t = time.perf_counter()
        gm_r, gl_r, _ = step()
        torch.cuda.synchronize()
        refined_ms = min(refined_ms, (time.perf_counter() - t) * 1e3)
    fwd_r, adj_r = dict(system.forward_stats), dict(system.adjoint_stats)
    g = torch.cat([gm, gl])
    g_r = torch.cat([gm_r, gl_r])
    rel_to_refined = lambda a: float(torch.linalg.vector_norm(a - g_r) / torch.linalg.vector_norm(g_r))
    err = rel_to_refined(g)
    err_mu = float(torch.linalg.vector_norm(gm - gm_r) / torch.linalg.vector_norm(gm_r))
    # the tolerance whose gradient meets 1e-4: both solves to 1e-5 (one full
    # FP32 solve and a short FP64-refined correction each)
    system.tol = 1e-5
    step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        gm5, gl5, _ = step()
    torch.cuda.synchronize()
    tight_ms = (time.perf_counter() - t) * 1e3 / reps
    fwd5, adj5 = dict(system.forward_stats), dict(system.adjoint_stats)
    err5 = rel_to_refined(torch.cat([gm5, gl5]))
    system.tol = args.tol
    out = {"workload": f"fem{n}: {n}^3 hex grid, {w['n_vertices']} vertices, {len(w['tets'])} tets, lognormal "
                       "Lame parameters, L = 1/2 ||x - x*||^2",
           "dof": 3 * w["n_vertices"], "upper_blocks": int(system.nb), "setup_ms": round(setup_ms, 1),
           "step_ms": round(step_ms, 3), "tol": args.tol,
           "assemble_ms": round(assemble_ms, 3), "assemble_and_build_ms": round(build_ms, 3),
           "one_solve_ms": round(solve_ms, 3),
           "grad_kernel_ms": round(grad_ms, 3),
           "forward": {"iterations": fwd["iterations"], "rel_res_true": fwd["rel_res_true"]},
           "adjoint": {"iterations": adj["iterations"], "rel_res_true": adj["rel_res_true"]},
           "refined": {"tol": 1e-9, "step_ms": round(refined_ms, 3),
                       "forward_iterations": fwd_r["iterations"], "forward_refinements": fwd_r["refinements"],
                       "adjoint_iterations": adj_r["iterations"], "adjoint_refinements": adj_r["refinements"]},
           "grad_rel_err_vs_refined": err, "grad_mu_rel_err_vs_refined": err_mu,
           "tol_1e-5": {"step_ms": round(tight_ms, 3), "forward_iterations": fwd5["iterations"],
                        "forward_refinements": fwd5["refinements"], "adjoint_iterations": adj5["iterations"],
                        "grad_rel_err_vs_refined": err5}}
    log(f"  adjoint: step {step_ms:.2f} ms (assemble {assemble_ms:.3f}, assemble + build {build_ms:.3f}, "
        f"one solve {solve_ms:.3f} (two per step), "
        f"grad {grad_ms:.3f}), iterations {fwd['iterations']}/{adj['iterations']}, refined step {refined_ms:.1f} ms, "
        f"grad rel err vs refined {err:.2e}; tol 1e-5: step {tight_ms:.2f} ms, grad rel err {err5:.2e}")
    return out


def fem40_leg(args, structure=True):
    """The 40^3 grid with 20,000 contacts (68,921 vertices, 206,763 DOF, 4 MAS
    levels): one system per step, plain FP32 vs FP16 / BF16 inverse storage,
    the roofline of its dominant kernel and (structure=True) the cost of a
    structure change (new contact set).  Returns the figures, the workload and
    the FP32 solution."""
    import numpy as np
    import torch
    from agipc_learning_amd import session
    from agipc_learning_amd import workloads as W

    log(f"step: generate {args.n}^3 grid, {args.contacts} contacts")
    w = W.fem_grid(args.n, young_sigma=0.5, n_contacts=args.contacts, seed=0)
    nv = w["n_vertices"]
    dof = 3 * nv
    log(f"step: analyze + setup (vertices {nv}, upper blocks {w['nb']})")
    t0 = time.perf_counter()
    s = session.solver_for(w, levels=args.levels)
    setup_ms = (time.perf_counter() - t0) * 1e3
    info = s.info()
    vals = torch.from_numpy(np.ascontiguousarray(w["vals"])).cuda()
    b = torch.from_numpy(session.rhs_of(w)).cuda()
    x = torch.zeros_like(b)
    bs = rhs_variants(b)
    torch.cuda.synchronize()
    work = torch.cuda.Stream()  # the step on one stream, as in the batch leg
    ws_ = work.cuda_stream

    def step(prec, k=0):
        s.set_values(0, vals, stream=ws_)
        s.build(prec, stream=ws_)
        return s.solve(bs[k % len(bs)], x, rule=1, tol=args.tol, stream=ws_)[0]

    def run(prec, steps, warmup):
        for i in range(warmup):
            step(prec, i + 1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = None
        for i in range(steps):
            st = step(prec, steps - 1 - i)  # the last step solves the original b
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / steps
        return ms, st

    s.build("fp32")
    if args.spmv_tiling:
        s.set_spmv_tiling(args.spmv_tiling)
        tiling = {"entries_per_thread": args.spmv_tiling, "ms": s.time_kernel(2, 50),
                  "nontemporal": s.spmv_nontemporal()}
    else:
        log("step: autotune SpMV tiling")
        tiling = s.autotune(50)
    log(f"  tiling {tiling}")
    precs = ["fp32"] + [p for p in args.precisions.split(",") if p and p != "fp32"]
    results = {}
    tiling["mas_local"] = {}
    for prec in precs:
        s.build(prec)
        tiling["mas_local"][prec] = s.autotune_mas(50)
        log(f"step: fem40c {prec} ({args.warmup} warmup + {args.fem40_steps} steps), local variant "
            f"{tiling['mas_local'][prec]}")
        ms, st = run(prec, args.fem40_steps, args.warmup)
        if st["failure"] != 0 or st["rel_res_true"] > args.tol * (1 + 1e-6):
            raise SystemExit(f"{prec} solve did not reach the tolerance: {st}")
        # split of the step: build alone, solve alone
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(5):
            s.build(prec)
        torch.cuda.synchronize()
        build_ms = (time.perf_counter() - t) * 1e3 / 5
        t = time.perf_counter()
        for _ in range(5):
            s.solve(b, x, rule=1, tol=args.tol)
        torch.cuda.synchronize()
        pcg_ms = (time.perf_counter() - t) * 1e3 / 5
        if prec == "fp32":
            x_fp32 = x.cpu().numpy().copy()
        kern = kernel_times(s, 100)
        s.build(prec)  # time_kernel leaves the iteration state frozen; rebuild
        results[prec] = dict(ms=ms, st=st, build_ms=build_ms, pcg_ms=pcg_ms, kernels_us=kern)
        log(f"  {prec}: {ms:.3f} ms/step (build {build_ms:.3f}, pcg {pcg_ms:.3f}), iterations {st['iterations']}, "
            f"true rel res {st['rel_res_true']:.2e}, kernels us {kern}")

    plain = results["fp32"]
    alg = algorithmic_bytes(info, nv, 4)
    kern = plain["kernels_us"]
    dominant = max(kern, key=lambda k: kern[k])
    achieved = alg[dominant] / (kern[dominant] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(dominant, (args.n, args.contacts, args.levels) == (40, 20000, 4),
                                            tiling["entries_per_thread"],
                                            waves=tiling["mas_local"]["fp32"]["waves"],
                                            spmv_nt=tiling["nontemporal"]),
                "algorithmic_bytes": alg[dominant], "avg_us": round(kern[dominant], 2),
                "per_kernel": {k: {"us": round(kern[k], 2), "GB/s": round(alg[k] / (kern[k] * 1e-6) / 1e9, 1)}
                               for k in kern}}

    if not structure:
        restructure = None
    else:
        # A Newton step whose contact set changed needs a new structure: the
        # analysis (hierarchy, block CSR, Galerkin maps, layouts) and device
        # set-up, then values and build.  Timed on the same mesh with another
        # contact set (mesh neighbours are per mesh): on a new solver, and on
        # the same solver (`ms`, the per-step path).
        log("step: structure change (new contact set)")
        from agipc_learning_amd import native

        w2 = W.fem_grid(args.n, young_sigma=0.5, n_contacts=args.contacts, seed=1)
        nbr = W.neighbors(w2["n_vertices"], w2["tets"])
        vals2 = torch.from_numpy(np.ascontiguousarray(w2["vals"])).cuda()
        torch.cuda.synchronize()
        def restructure_once():
            t0 = time.perf_counter()
            s2 = native.Solver()
            s2.add_system(block_rows=w2["block_rows"], rows=w2["rows"], cols=w2["cols"], use_mas=True,
                          levels=args.levels)
            s2.set_topology_morton(0, *nbr, w2["positions"])
            s2.set_contacts(0, w2["pairs"])
            t1 = time.perf_counter()
            s2.setup()  # MAS hierarchy on the device, host plan, device upload
            t2 = time.perf_counter()
            s2.set_values(0, vals2)
            s2.build("fp32")
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            return (t3 - t0) * 1e3, (t2 - t1) * 1e3, (t3 - t2) * 1e3

        restructure_once()
        times = sorted((restructure_once() for _ in range(3)), key=lambda r: r[0])
        tot, setup_ms, build_ms = times[1]
        # the host-only analysis (hierarchy built on the host), for comparison
        s3 = native.Solver()
        s3.add_system(block_rows=w2["block_rows"], rows=w2["rows"], cols=w2["cols"], use_mas=True, levels=args.levels)
        s3.set_topology_morton(0, *nbr, w2["positions"])
        s3.set_contacts(0, w2["pairs"])
        t = time.perf_counter()
        s3.analyze()
        host_analysis_ms = (time.perf_counter() - t) * 1e3
        restructure = {"new_solver_ms": round(tot, 2), "new_solver_setup_ms": round(setup_ms, 2),
                       "values_and_build_ms": round(build_ms, 2),
                       "host_only_analysis_ms": round(host_analysis_ms, 2)}
        del s3
        # the same solver taking each new structure (masx_set_structure + contacts
        # + setup + values + build): mesh topology kept, host and device buffers
        # reused; the two contact sets alternate so every call is a real change
        # (block CSR, Galerkin maps, bank entries and SpMV layout analyzed on the
        # device, the default for one MAS system; and, for comparison, on the host)
        for key, on_device in (("same_solver_ms", True), ("same_solver_host_structure_ms", False)):
            s4 = native.Solver()
            s4.add_system(block_rows=w["block_rows"], rows=w["rows"], cols=w["cols"], use_mas=True,
                          levels=args.levels)
            s4.set_device_structure(on_device)
            s4.set_topology_morton(0, *nbr, w2["positions"])
            s4.set_contacts(0, w["pairs"])
            s4.setup()
            same = []
            for i in range(8):
                ww, vv = (w2, vals2) if i % 2 == 0 else (w, vals)
                torch.cuda.synchronize()
                t = time.perf_counter()
                s4.set_structure(0, ww["rows"], ww["cols"])
                s4.set_contacts(0, ww["pairs"])
                s4.setup()
                s4.set_values(0, vv)
                s4.build("fp32")
                torch.cuda.synchronize()
                same.append((time.perf_counter() - t) * 1e3)
            if s4.device_structure_used() != on_device:
                raise SystemExit("structure analysis ran on the wrong side")
            restructure[key] = round(sorted(same[2:])[2], 2)
            s4.close()
        # the per-Newton-step path (the solver kept, as the reference keeps its
        # MAS object across setPreconditioner_bcoo calls)
        restructure = {"ms": restructure["same_solver_ms"], **restructure}
        del vals2
        # the MAS hierarchy alone (the reference's per-step ReorderRealtime): host
        # build vs the device build, inputs resident where each one runs
        p2r, r2p = native.morton_partition(w2["positions"])
        n2, slots2 = w2["n_vertices"], len(p2r)
        host_ms = []
        for _ in range(3):
            t = time.perf_counter()
            hh = native.hierarchy_host(n2, slots2, *nbr, p2r, r2p, pairs=w2["pairs"])
            host_ms.append((time.perf_counter() - t) * 1e3)
        cu = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).cuda()
        d_in = [cu(nbr[0], np.uint32), cu(nbr[1], np.uint32), cu(nbr[2], np.uint32), cu(p2r, np.int32),
                cu(r2p, np.int32)]
        d_pairs = cu(np.asarray(w2["pairs"]).reshape(-1, 4), np.int32)
        hier = native.DeviceHierarchy()
        dev_ms = []
        for _ in range(6):
            torch.cuda.synchronize()
            t = time.perf_counter()
            hier.build(n2, slots2, *d_in, pairs=d_pairs)
            dev_ms.append((time.perf_counter() - t) * 1e3)
        dh = hier.download()
        same = all(np.array_equal(dh[k], hh[k]) for k in ("fine_mask", "prefix", "coarse_table", "level_begin"))
        if not same:
            raise SystemExit("device hierarchy differs from the host build")
        restructure["hierarchy_host_ms"] = round(sorted(host_ms)[1], 3)
        restructure["hierarchy_device_ms"] = round(sorted(dev_ms[1:])[2], 3)
        # the Morton partition the banks are cut from: host vs device (same arrays)
        mt_host, mt_dev = [], []
        dm = native.DeviceMorton()
        d_pos = torch.from_numpy(np.ascontiguousarray(w["positions"], np.float64)).cuda()
        for _ in range(5):
            t = time.perf_counter()
            native.morton_partition(w["positions"], 32)
            mt_host.append((time.perf_counter() - t) * 1e3)
            torch.cuda.synchronize()
            t = time.perf_counter()
            dm.partition(d_pos, 32)
            torch.cuda.synchronize()
            mt_dev.append((time.perf_counter() - t) * 1e3)
        restructure["morton_host_ms"] = round(sorted(mt_host)[2], 3)
        restructure["morton_device_ms"] = round(sorted(mt_dev[1:])[1], 3)
        dm.close()
        log(f"  structure change: {restructure}")
        del hier, d_in, d_pairs

    iters = plain["st"]["iterations"]
    out = {"workload": f"fem{args.n}c: {args.n}^3 hex grid, {nv} vertices, {dof} DOF, {len(w['tets'])} tets, "
                       f"{args.contacts} contact pairs, lognormal E, 32-vertex Morton banks, {args.levels} MAS levels",
           "vertices": nv, "dof": dof, "upper_blocks": int(info["nb_total"]), "banks": int(info["n_banks"]),
           "ms": round(plain["ms"], 4), "iterations": iters, "rel_res_true": plain["st"]["rel_res_true"],
           "solves_per_s": round(1e3 / plain["ms"], 2),
           "dof_iterations_per_s": round(dof * iters / (plain["pcg_ms"] * 1e-3), 1),
           "plain_fp32": {"ms": round(plain["ms"], 4), "build_ms": round(plain["build_ms"], 4),
                          "pcg_ms": round(plain["pcg_ms"], 4), "iterations": iters},
           "mixed": {p: {"ms": round(results[p]["ms"], 4), "build_ms": round(results[p]["build_ms"], 4),
                         "pcg_ms": round(results[p]["pcg_ms"], 4), "iterations": results[p]["st"]["iterations"],
                         "rel_res_true": results[p]["st"]["rel_res_true"],
                         "speedup_vs_fp32": round(plain["ms"] / results[p]["ms"], 3)} for p in precs[1:]},
           "setup_ms": round(setup_ms, 1), "spmv_tiling": tiling, "roofline": roofline}
    if restructure is not None:
        out["structure_change"] = restructure
    return out, w, x_fp32


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    import torch

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_leg(args)}), flush=True)
        return
    if args.fem40_only:
        print(json.dumps({"fem40c": fem40_leg(args, structure=False)[0]}), flush=True)
        return
    if args.hybrid_only:
        from agipc_learning_amd import workloads as W

        print(json.dumps({"hybrid": hybrid(args, W.hybrid(seed=0))[0]}), flush=True)
        return

    batch, ws, batch_sols = batch_leg(args)
    if args.batch_only:
        print(json.dumps({"batch64": batch}), flush=True)
        return
    fem40 = hyb = adj = None
    x40 = w40 = hw = hyb_sols = None
    if not args.no_extra:
        from agipc_learning_amd import workloads as W

        fem40, w40, x40 = fem40_leg(args)
        hw = W.hybrid(seed=0)
        hyb, hyb_sols = hybrid(args, hw)
        adj = adjoint_leg(args)
    log("step: streaming rates")
    ceil = stream_ceilings()
    log(f"  {ceil}")
    rf = batch["roofline"]
    rf["read_stream_GBs"], rf["copy_GBs"] = ceil["read_GBs"], ceil["copy_GBs"]
    rf["frac_of_read_stream"] = round(rf["achieved"] / ceil["read_GBs"], 4)

    cpu = None
    if not args.no_cpu:
        O = oracle_module()
        log(f"step: cpu baseline, 64-system batch ({O.get_threads()} threads, {args.cpu_runs} runs; one thread on "
            "a sample)")
        cpu, sols, oracle_its = cpu_batch(O, ws, args.tol, args.cpu_runs)
        log(f"  cpu batch: {cpu['all_cores']} / one thread {cpu['single_core']}")
        for prec, xg in batch_sols.items():
            off, errs = 0, []
            for w, xo in zip(ws, sols):
                n = 3 * w["block_rows"]
                errs.append(rel_err(xg[off:off + n], xo))
                off += n
            gi = batch[prec]["iterations"]
            batch[prec]["solution_rel_err_vs_oracle"] = {"max": max(errs), "median": sorted(errs)[len(errs) // 2]}
            batch[prec]["iterations_minus_oracle"] = [min(a - b for a, b in zip(gi, oracle_its)),
                                                     max(a - b for a, b in zip(gi, oracle_its))]
        batch["oracle_iterations"] = {"max": max(oracle_its), "mean": round(sum(oracle_its) / len(oracle_its), 2)}
        del sols
        if fem40 is not None:
            log("step: cpu baseline, 40^3 + contacts")
            c40, xo = cpu_single(O, w40, args.levels, args.tol, max(args.cpu_runs, 5))
            fem40["cpu_baseline"] = c40
            fem40["solution_rel_err_vs_oracle"] = rel_err(x40, xo)
            log(f"  cpu 40^3: {c40['all_cores']} / one thread {c40['single_core']}, solution rel err "
                f"{fem40['solution_rel_err_vs_oracle']:.2e}")
        if hyb is not None:
            log("step: oracle, hybrid")
            ch, xo = cpu_single(O, hw, 0, args.tol, args.cpu_runs)
            hyb["cpu_baseline"] = ch
            hyb["oracle_iterations"] = ch["oracle_iterations"]
            for prec, xg in hyb_sols.items():
                hyb[prec]["solution_rel_err_vs_oracle"] = rel_err(xg, xo)
            log(f"  hybrid oracle: {ch['oracle_iterations']} iterations, errors "
                f"{[hyb[p]['solution_rel_err_vs_oracle'] for p in hyb_sols]}")

    plain = batch["fp32"]
    out = {
        "metric": "MAS-PCG solve ms to 1e-4 rel-res, mixed vs plain FP32; % HBM roofline",
        "value": plain["ms"],
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batch64: 64 independent FEM + contact systems in one batched solve, cubes of "
                               f"U[5k, 80k] vertices (seed 0), {batch['vertices']} vertices, {batch['dof']} DOF, "
                               f"{batch['upper_blocks']} upper 3x3 blocks, lognormal E, 5% of vertices in contact "
                               "pairs, 32-vertex Morton banks, tol 1e-4 (FP64 true residual) per system, step = "
                               "values in + MAS build + PCG of all 64",
                   "systems": batch["systems"], "vertices": batch["vertices"], "dof": batch["dof"],
                   "upper_blocks": batch["upper_blocks"], "banks": batch["banks"]},
        "iterations": plain["max_iterations"],
        "rel_res_true": plain["worst_rel_res_true"],
        "solves_per_s": round(1e3 / plain["ms"] * batch["systems"], 2),
        "dof_iterations_per_s": plain["dof_iterations_per_s"],
        "plain_fp32": {k: plain[k] for k in ("ms", "build_ms", "pcg_ms", "max_iterations", "mean_iterations")},
        "mixed": {p: {k: batch[p][k] for k in ("ms", "build_ms", "pcg_ms", "max_iterations", "mean_iterations",
                                               "speedup_vs_fp32", "worst_rel_res_true")}
                  for p in batch if p in ("fp16", "bf16")},
        "roofline": rf,
        "cpu_baseline": cpu,
        "stream": ceil,
        "batch64": batch,
        "fem40c": fem40,
        "hybrid": hyb,
        "adjoint": adj,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
