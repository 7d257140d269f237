"""A/B of library builds (or solve settings) in one process on one GPU.

  python -m agipc_learning_amd.ab_compare LIB[:CHUNK[:TILING]] LIB[:CHUNK[:TILING]] ... [--rounds R]
      [--precision fp32|fp16|bf16] [--workload fem40c|batch64]

Each argument is a built libmasx.so, optionally with the solve's graph chunk
(iterations per captured graph, default 8) and the SpMV tiling (entries per
thread, default 4).  Every variant sets up the workload (default: the 40^3
grid with 20,000 contacts, 4 MAS levels; or the 64-system batch) and then,
in alternating rounds, times five steps (values in, build in the chosen
inverse precision, PCG to 1e-4; one step for the batch) and its four
iteration kernels.  Medians over the
rounds are printed per variant: on these machines a whole run can be a few
per cent faster or slower than another, so two builds are compared inside
one process, interleaved, rather than across runs.
"""
from __future__ import annotations

import argparse
import os
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("variants", nargs="+")
    ap.add_argument("--rounds", type=int, default=12)
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp16", "bf16"))
    ap.add_argument("--workload", default="fem40c", choices=("fem40c", "batch64"))
    args = ap.parse_args(argv)
    import numpy as np
    import torch

    from . import native, session
    from . import workloads as W

    specs = []
    for v in args.variants:
        parts = v.split(":")
        specs.append((os.path.abspath(parts[0]), int(parts[1]) if len(parts) > 1 else 8,
                      int(parts[2]) if len(parts) > 2 else 4))
    prec = args.precision
    if args.workload == "batch64":
        ws = W.batch64(seed=0)
        n_steps = 1
    else:
        ws = [W.fem_grid(40, young_sigma=0.5, n_contacts=20000, seed=0)]
        n_steps = 5
    vals = [torch.from_numpy(np.ascontiguousarray(w["vals"])).cuda() for w in ws]
    b = torch.from_numpy(session.rhs_of(ws)).cuda()
    solvers = []
    for lib, _, tiling in specs:
        native.LIB_PATH = lib
        native._lib = None  # each Solver keeps the library it was created with
        s = session.solver_for(ws, levels=4 if len(ws) == 1 else 0)
        s.build(prec)
        s.set_spmv_tiling(tiling)
        solvers.append(s)
    xs = [torch.zeros_like(b) for _ in solvers]

    def step(i):
        s = solvers[i]
        for k, v in enumerate(vals):
            s.set_values(k, v)
        s.build(prec)
        return s.solve(b, xs[i], rule=1, tol=1e-4, chunk=specs[i][1])

    for i in range(len(solvers)):
        for _ in range(3):
            step(i)
    torch.cuda.synchronize()
    steps = {i: [] for i in range(len(solvers))}
    kern = {i: [[] for _ in range(4)] for i in range(len(solvers))}
    last = {}
    for r in range(args.rounds):
        for i, s in enumerate(solvers):
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(n_steps):
                st = step(i)
            torch.cuda.synchronize()
            steps[i].append((time.perf_counter() - t) / n_steps * 1e3)
            last[i] = (max(x["iterations"] for x in st), max(x["rel_res_true"] for x in st))
            for k in range(4):
                kern[i][k].append(s.time_kernel(k, 50 if n_steps > 1 else 20) * 1e3)
            s.build(prec)  # time_kernel leaves the iteration state frozen
        print(f"round {r} done", flush=True)
    for i, (lib, chunk, tiling) in enumerate(specs):
        km = [float(np.median(v)) for v in kern[i]]
        print(f"{os.path.basename(lib)}:{chunk}:{tiling} {prec} step median {np.median(steps[i]):.4f} ms "
              f"(p25 {np.percentile(steps[i], 25):.4f} p75 {np.percentile(steps[i], 75):.4f}), iterations "
              f"{last[i][0]}, true rel res {last[i][1]:.3e} | kernels us: MAS local {km[0]:.2f}, finish {km[1]:.2f}, "
              f"SpMV {km[2]:.2f}, update {km[3]:.2f}")


if __name__ == "__main__":
    main()
