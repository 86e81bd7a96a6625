#!/usr/bin/env python3
"""Benchmark: Sz=0 H.v GB/s and Lanczos steps/s on the 4x8 periodic square
lattice (N=32, J1=D1=1, J2=D2=0.5, D = 601,080,390), one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one three-term Lanczos step (H.v with the alpha dot fused, then
the fused update + norm).  `value` is the algorithmic H.v bandwidth
B_Hv / t_Hv with B_Hv = 8 (2 D + nnz_off), nnz_off = n_bonds * 2 C(N-2, N/2-1)
(one 8 B gather per off-diagonal nonzero, the x_r read and the y_r write),
timed with HIP events on the launch stream over the timed region.  Inputs
are resident in HBM before timing starts.  The CPU baseline is the oracle's
matrix-free H.v (same rank arithmetic, OpenMP over the host cores) on a
bounded contiguous row sample of the same problem.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Sz=0 H·v GB/s + Lanczos steps/s at N=32; batched realizations/s vs plain"


def algorithmic_bytes(N, nup, n_bonds_flip):
    D = math.comb(N, nup)
    nnz = n_bonds_flip * 2 * math.comb(N - 2, nup - 1)
    return 8 * (2 * D + nnz), D, nnz


def measured_traffic(kernel_name, lx, ly):
    """HBM bytes per H.v launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, calibrated against k_lanczos_update's known
    traffic); null when no measurement of this kernel and lattice exists."""
    path = os.path.join(ROOT, "profiles", f"hv_traffic_{lx}x{ly}.json")
    try:
        with open(path) as fh:
            return json.load(fh)["kernels"][kernel_name]["hbm_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(h_args, target_s=12.0):
    """Oracle matrix-free H.v rows/s on this host, scaled to GB/s of B_Hv."""
    import numpy as np
    from oracle import oracle as O

    N, nup, J1, D1, J2, D2, lx, ly, B = h_args
    m = O.Model(O.SQUARE, N, nup, J1, D1, J2, D2, lx=lx, ly=ly)
    D = m.D
    x = np.random.default_rng(0).random(D)
    threads = O.lib().oq_max_threads()
    n = 200_000
    t = time.perf_counter()
    m.hv_matfree(x, row0=0, nrows=n, nthreads=threads)
    dt = time.perf_counter() - t
    rate = n / dt
    n2 = int(min(D, max(n, rate * target_s)))
    t = time.perf_counter()
    m.hv_matfree(x, row0=D // 3, nrows=n2, nthreads=threads)
    dt = time.perf_counter() - t
    rows_per_s = n2 / dt
    gbs = B / D * rows_per_s / 1e9
    return {"value": round(gbs, 4), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"oracle matrix-free H.v (oracle/oracle.cpp oq_hv_matfree, exact binomial ranks) over rows "
                      f"[{D // 3}, {D // 3 + n2}) of the same 4x8 problem ({n2} rows, {dt:.1f} s, {threads} threads); "
                      f"GB/s = B_Hv/D x rows/s; the reference binary (PETSc/SLEPc) cannot be built here",
            "rows_per_s": round(rows_per_s, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lx", type=int, default=4)
    ap.add_argument("--ly", type=int, default=8)
    ap.add_argument("--hv-kernel", type=int, default=0, help="0 auto, 1 generic, 2 row split, 3 specialised row split")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the Sz-sector path here is built and measured on a single "
              "MI355X (multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)

    import torch

    import lsquantumed_amd as Q

    torch.cuda.set_device(0)
    lx, ly = args.lx, args.ly
    N = lx * ly
    J1 = D1 = 1.0
    J2 = D2 = 0.5
    env = Q.Environment(N)
    basis = Q.Basis(env, 0)
    lat = Q.square2D(env, lx, ly)
    h = Q.Hamiltonian(env, basis, lat, J1, D1, J2, D2)
    h.build_diag(env)
    h.build_off_diag(env)
    if args.hv_kernel:
        h.set_hv_kernel(args.hv_kernel)
    kernel_name = {1: "generic", 2: "row split", 3: "row split (hipRTC)"}[h.hv_kernel()]
    nflip = len(lat.nn_bonds[0]) + len(lat.nnn_bonds[0])
    B, D, nnz = algorithmic_bytes(N, N // 2, nflip)
    assert D == h.hamilt_size()

    g = torch.Generator(device="cuda").manual_seed(1)
    q = torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5
    p = torch.zeros(D, dtype=torch.float64, device="cuda")
    y = torch.empty(D, dtype=torch.float64, device="cuda")
    sq = 1.0 / torch.linalg.vector_norm(q).item()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    # --- H.v alone (the dominant kernel) ----------------------------------
    for _ in range(args.warmup):
        h.matvec(q, y)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        h.matvec(q, y)
    ev1.record(stream)
    torch.cuda.synchronize()
    hv_ms = ev0.elapsed_time(ev1) / args.steps
    hv_gbs = B / (hv_ms * 1e-3) / 1e9

    # --- Lanczos steps ------------------------------------------------------
    sp, beta = 0.0, 0.0

    def step():
        nonlocal q, p, sq, sp, beta
        a, bn = Q.lanczos_step(h, q, sq, p, sp, beta, y)
        q, p = p, q
        sp, sq, beta = sq, 1.0 / bn, bn
        return a

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    alphas = [step() for _ in range(args.steps)]
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = ev0.elapsed_time(ev1) / args.steps
    steps_per_s = 1e3 / step_ms

    cpu = None if args.no_cpu else cpu_baseline((N, N // 2, J1, D1, J2, D2, lx, ly, B))
    out = {
        "metric": METRIC,
        "value": round(hv_gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform start vector)",
        "config": {"workload": f"Sz=0 H.v + 3-term Lanczos step, periodic square {lx}x{ly} (N={N}), "
                               f"J1=D1=1, J2=D2=0.5, D={D}", "D": D, "nnz_offdiag": nnz,
                   "algorithmic_bytes_per_hv": B, "parallelism": "single GPU",
                   "hv_kernel": kernel_name},
        "hv_ms": round(hv_ms, 3),
        "lanczos_steps_per_s": round(steps_per_s, 3),
        "lanczos_wall_s": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(hv_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(hv_gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel_name, lx, ly)},
        "cpu_baseline": cpu,
        "alpha_last": alphas[-1],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
