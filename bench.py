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


def log(msg):
    """Progress line (stderr: stdout carries the one JSON result line)."""
    print(msg, file=sys.stderr, flush=True)
METRIC = "Sz=0 H·v GB/s + Lanczos steps/s at N=32; batched realizations/s vs plain"


def algorithmic_bytes(N, nup, n_bonds_flip):
    D = math.comb(N, nup)
    nnz = n_bonds_flip * 2 * math.comb(N - 2, nup - 1)
    return 8 * (2 * D + nnz), D, nnz


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hv_traffic.json")

# HBM bytes of one dispatch from the L2 (TCC) -> memory request counters of
# gfx950: reads are 128-, 64- or 32-byte requests (TCC_EA0_RDREQ_{128B,64B,32B}),
# writes are counted in 32-byte units going to DRAM (TCC_EA0_WRREQ_WRITE_DRAM_32B).
# rocprofv3's FETCH_SIZE counts 128-byte requests through TCC_BUBBLE, which
# reads 0 on gfx950, so it reports half of a streaming read; the request
# counters above are exact on the probes of `--calibrate` (read stream, triad,
# and one double per 128-byte line or per 64-byte half line: both cost one
# 128-byte request per access).
READ_COUNTERS = {"TCC_EA0_RDREQ_128B_sum": 128, "TCC_EA0_RDREQ_64B_sum": 64, "TCC_EA0_RDREQ_32B_sum": 32}
WRITE_COUNTERS = {"TCC_EA0_WRREQ_WRITE_DRAM_32B_sum": 32}


def measured_traffic(fingerprint):
    """The counter record (HBM bytes of one launch) of the H.v whose
    fingerprint (lsqed_ham_hv_fingerprint: kernel, grid, model and the hash
    of the generated kernel source) is `fingerprint`; None when no record of
    exactly this kernel exists, so a changed kernel never inherits old bytes."""
    try:
        with open(TRAFFIC_FILE) as fh:
            return json.load(fh)["records"].get(fingerprint)
    except (OSError, KeyError, ValueError):
        return None


def record_traffic(prefix, kernel="hv_jit_nodot"):
    """Build the traffic record of one H.v from rocprofv3 --pmc passes
    `<prefix>_<i>/run_counter_collection.csv` (made with `--hv-only`, whose
    JSON line in `<prefix>_<i>.log` names the fingerprint) and store it in
    profiles/hv_traffic.json under that fingerprint."""
    import csv

    vals, fp, hv = {}, None, None
    for i in range(1, 33):
        d = f"{prefix}_{i}"
        f = os.path.join(d, "run_counter_collection.csv")
        if not os.path.exists(f):
            continue
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if kernel in row["Kernel_Name"]:
                    vals.setdefault(row["Counter_Name"], []).append(float(row["Counter_Value"]))
        if os.path.exists(d + ".log"):
            for ln in open(d + ".log"):
                if ln.startswith('{"hv"'):
                    hv = json.loads(ln)["hv"]
                    fp = fp or hv["fingerprint"]
                    if hv["fingerprint"] != fp:
                        raise SystemExit(f"record_traffic: {d} profiled another kernel")
    mean = {k: sum(v) / len(v) for k, v in vals.items()}
    missing = [c for c in list(READ_COUNTERS) + list(WRITE_COUNTERS) if c not in mean]
    if missing or not fp:
        raise SystemExit(f"record_traffic: missing counters {missing} or fingerprint under {prefix}")
    rd = sum(mean[c] * w for c, w in READ_COUNTERS.items())
    wr = sum(mean[c] * w for c, w in WRITE_COUNTERS.items())
    try:
        with open(TRAFFIC_FILE) as fh:
            doc = json.load(fh)
    except (OSError, ValueError):
        doc = {}
    doc.setdefault("method", "HBM bytes per dispatch = 128 TCC_EA0_RDREQ_128B + 64 TCC_EA0_RDREQ_64B + "
                             "32 TCC_EA0_RDREQ_32B (reads) + 32 TCC_EA0_WRREQ_WRITE_DRAM_32B (writes), "
                             "mean over the profiled dispatches of the kernel")
    doc.setdefault("records", {})[fp] = {
        "workload": hv["workload"], "kernel": kernel, "hv_kernel": hv["hv_kernel"],
        "read_bytes": int(rd), "write_bytes": int(wr), "hbm_bytes": int(rd + wr),
        "counters": {k: round(v, 1) for k, v in sorted(mean.items())},
        "dispatches": max(len(v) for v in vals.values()),
        "profiled_hv_ms": hv["hv_ms"]}
    with open(TRAFFIC_FILE, "w") as fh:
        json.dump(doc, fh, indent=1, sort_keys=True)
        fh.write("\n")
    return fp, doc["records"][fp]


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


def disorder_batch(reps: int = 3):
    """Batched disorder ground states against the plain one-at-a-time path.

    Workload: 64 periodic chains of N=18, 64 of N=20 and 128 of N=22 (Sz=0
    dims 48,620 / 184,756 / 705,432), J1=D1=1, J2=D2=0.2, fields
    h_i ~ U[-2, 2] from mt19937_64(1 + k) for the k-th realisation, Lanczos to
    a 1e-9 relative residual.  Batched: one BatchSolver over all 256 (warm-up
    solve = autotune + kernel compilation, then `reps` timed solves).  Plain:
    the realisation loop of examples/main_disorder.cpp:82-120 on the GPU --
    one model per N, build_disorder, SolverInit(nev=1), solve, solution,
    SolverClean per realisation (Krylov-Schur, tol 1e-9), after one untimed
    solve per N, with the refined CGS2 of SLEPc's Krylov-Schur (the
    reference's solver); also with the library's default orthogonalisation
    (delayed CGS2 on sectors of 2^16 states or more).  Also timed: the
    batched solver fed one realisation at a time (batch size 1)."""
    import numpy as np
    import torch

    import lsquantumed_amd as Q
    from lsquantumed_amd.batch import BatchSolver, disorder_chains

    sizes, counts = [18, 20, 22], [64, 64, 128]
    hams = disorder_chains(sizes, counts, J1=1.0, J2=0.2, W=2.0, seed=1)
    n = len(hams)
    b = BatchSolver(hams, tol=1e-9)
    t = time.perf_counter()
    b.solve()
    first = time.perf_counter() - t
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        e0 = b.solve()
        times.append(time.perf_counter() - t)
    t_b = float(np.mean(times))
    info = [b.info(k) for k in range(n)]
    sched = b.schedule()
    b.destroy()

    # plain: one model per N reused, fields rebuilt per realisation
    def plain_loop(delayed):
        saved = os.environ.get("LSQED_KS_DELAYED")
        if delayed is not None:
            os.environ["LSQED_KS_DELAYED"] = "1" if delayed else "0"
        try:
            per_n = {}
            for N in sizes:
                env = Q.Environment(N)
                h = Q.Hamiltonian(env, Q.Basis(env, 0), Q.chain1D(env), 1.0, 1.0, 0.2, 0.2)
                h.build_diag(env)
                h.build_off_diag(env)
                h.build_disorder(env, np.zeros(N))
                eps = Q.Solver.SolverInit(h, 1, tol=1e-9)
                Q.Solver.solve(env, eps)
                Q.Solver.SolverClean(eps)
                per_n[N] = (env, h)
            e = np.zeros(n)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for k, hk in enumerate(hams):
                env, h = per_n[hk.basis.nspins]
                h.build_disorder(env, hk.hi)
                eps = Q.Solver.SolverInit(h, 1, tol=1e-9)
                if Q.Solver.solve(env, eps) < 1:
                    raise RuntimeError(f"plain solve of realisation {k} did not converge")
                e[k], _ = Q.Solver.solution(eps, 0)
                Q.Solver.SolverClean(eps)
            return e, time.perf_counter() - t
        finally:
            if saved is None:
                os.environ.pop("LSQED_KS_DELAYED", None)
            else:
                os.environ["LSQED_KS_DELAYED"] = saved

    plain, t_p = plain_loop(False)        # SLEPc's refined CGS2
    plain_d, t_pd = plain_loop(None)      # the library's default

    # the batched solver one realisation at a time
    singles = [BatchSolver([h], tol=1e-9) for h in hams]
    for N in sizes:  # warm-up per lattice (tuning, compilation)
        next(s for s, h in zip(singles, hams) if h.basis.nspins == N).solve()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for s in singles:
        s.solve()
    t_1 = time.perf_counter() - t
    for s in singles:
        s.destroy()

    rel = np.abs(e0 - plain) / np.abs(plain)
    rel_d = np.abs(e0 - plain_d) / np.abs(plain_d)
    # algorithmic bytes of one batched step with every model active: H.v
    # (x read, y written, one 8 B gather per off-diagonal nonzero) + update
    # (y, q, p read, p written)
    hv_bytes = sum(algorithmic_bytes(N, N // 2, 2 * N)[0] * c for N, c in zip(sizes, counts))
    upd_bytes = 32 * sum(math.comb(N, N // 2) * c for N, c in zip(sizes, counts))
    return {
        "workload": "64x N=18 + 64x N=20 + 128x N=22 periodic chains, J1=D1=1, J2=D2=0.2, h_i~U[-2,2] from "
                    "mt19937_64(1+k), Sz=0, ground state to 1e-9 relative residual",
        "realisations": n,
        "batched_realisations_per_s": round(n / t_b, 2),
        "batched_solve_s": round(t_b, 4),
        "batched_first_solve_s": round(first, 3),
        "lanczos_steps_mean": round(float(np.mean([i[0] for i in info])), 1),
        "all_converged": bool(all(i[2] for i in info)),
        "schedule": sched,
        "step_algorithmic_bytes": hv_bytes + upd_bytes,
        # effective bandwidth: one 8-byte read per gather, cache hits included (not an HBM fraction)
        "step_effective_gbs": round((hv_bytes + upd_bytes) / (sched["step_ms"] * 1e-3) / 1e9, 1)
        if sched["step_ms"] else None,
        "plain_realisations_per_s": round(n / t_p, 2),
        "plain_s": round(t_p, 3),
        "plain": "Krylov-Schur nev=1 per realisation (SolverInit/solve/solution/SolverClean, tol 1e-9) with "
                 "SLEPc's refined CGS2, the reference's solver",
        "plain_library_default_realisations_per_s": round(n / t_pd, 2),
        "plain_library_default": "the same loop with the library's default orthogonalisation (delayed CGS2 on "
                                 "sectors of 2^16 states or more: N=20, 22)",
        "batch_of_one_realisations_per_s": round(n / t_1, 2),
        "speedup_vs_plain": round(t_p / t_b, 2),
        "speedup_vs_plain_library_default": round(t_pd / t_b, 2),
        "max_rel_e0_diff_vs_plain": float(max(rel.max(), rel_d.max())),
    }


def stream_ceiling(q, p, y, reps: int = 10):
    """Measured streaming ceilings on the same buffers (HIP events): the
    library's read-only stream (8 B per element) and STREAM triad (24 B per
    element), and torch's elementwise add as a second triad."""
    import ctypes as C

    import torch

    from lsquantumed_amd._lib import check, lib

    n = q.numel() - (q.numel() & 1)
    part = torch.empty(2048, dtype=torch.float64, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(f):
        f()
        ev0.record()
        for _ in range(reps):
            f()
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / reps

    ms_r = timed(lambda: check(lib().lsqed_stream_probe(1, None, C.c_void_p(q.data_ptr()), None, n,
                                                         C.c_void_p(part.data_ptr()), st)))
    ms_t = timed(lambda: check(lib().lsqed_stream_probe(0, C.c_void_p(y.data_ptr()), C.c_void_p(q.data_ptr()),
                                                         C.c_void_p(p.data_ptr()), n, None, st)))
    ms_tt = timed(lambda: torch.add(q[:n], p[:n], alpha=0.5, out=y[:n]))
    return {"read_gbs": round(8 * n / ms_r / 1e6, 1), "triad_gbs": round(24 * n / ms_t / 1e6, 1),
            "torch_triad_gbs": round(24 * n / ms_tt / 1e6, 1), "elements": n}


def stream_sizes(reps: int = 20):
    """Read bandwidth of the library's read-only stream (lsqed_stream_probe
    kind 1) against the size of the buffer it streams, 16 MB to 4 GB: three
    warm-up passes, then `reps` passes over the same buffer timed with HIP
    events.  Buffers the memory-side cache (256 MB) or the L2s (8 x 4 MB)
    hold stream from there on repeated passes."""
    import ctypes as C

    import torch

    from lsquantumed_amd._lib import check, lib

    part = torch.empty(2048, dtype=torch.float64, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {}
    for mb in (16, 32, 64, 128, 192, 256, 384, 512, 1024, 4096):
        n = mb * (1 << 20) // 8
        b = torch.rand(n, dtype=torch.float64, device="cuda")

        def run():
            check(lib().lsqed_stream_probe(1, None, C.c_void_p(b.data_ptr()), None, n,
                                           C.c_void_p(part.data_ptr()), st))

        for _ in range(3):
            run()
        ev0.record()
        for _ in range(reps):
            run()
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        out[f"{mb} MB"] = round(8 * n / ms / 1e9, 2)  # TB/s
        del b
    torch.cuda.empty_cache()
    return {"unit": "TB/s", "passes": reps, "read_tbs": out}


def calibrate(reps: int = 2):
    """Kernels of known traffic for reading the HBM byte counters (rocprofv3
    --pmc): the read-only stream (8 B per element read), the STREAM triad
    (16 B read + 8 B written per element), and the scattered-line probes
    (one double per 128-byte line, resp. per 64-byte half line, every line
    touched once in an order with no locality).  Each is launched `reps`
    times on 2 GiB buffers; prints one JSON object with the algorithmic bytes
    of one launch of each."""
    import ctypes as C

    import torch

    from lsquantumed_amd._lib import check, lib

    n = 1 << 28
    b = torch.rand(n, dtype=torch.float64, device="cuda")
    c = torch.rand(n, dtype=torch.float64, device="cuda")
    a = torch.empty(n, dtype=torch.float64, device="cuda")
    part = torch.empty(2048, dtype=torch.float64, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    lines, halves = n // 16, n // 8
    for _ in range(reps):
        check(lib().lsqed_stream_probe(1, None, P(b), None, n, P(part), st))
        check(lib().lsqed_stream_probe(0, P(a), P(b), P(c), n, None, st))
        check(lib().lsqed_stream_probe(2, P(a), P(b), None, lines, None, st))
        check(lib().lsqed_stream_probe(3, P(a), P(b), None, halves, None, st))
    torch.cuda.synchronize()
    return {"k_stream_read": {"read": 8 * n, "write": 0},
            "k_stream_triad": {"read": 16 * n, "write": 8 * n},
            "k_stream_gather sh=4 (one double per 128-B line)": {"lines": lines, "useful_read": 8 * lines,
                                                                 "write": 8 * lines},
            "k_stream_gather sh=3 (one double per 64-B half line)": {"halves": halves, "useful_read": 8 * halves,
                                                                     "write": 8 * halves},
            "launches_each": reps}


def cpu_disorder_baseline():
    """CPU realisations/s for the 256-chain disorder workload: the reference's
    data layout (assembled CSR, complex128 values + int64 columns, as PETSc
    with complex scalars stores it) and a three-term Lanczos on complex
    vectors to the same 1e-9 relative residual, OpenMP SpMV over the host
    cores; timed on a sample of realisations per N (the assembly of each
    lattice's matrix is timed once and reported apart, since the reference
    reuses the off-diagonal part across realisations)."""
    import numpy as np

    from oracle import oracle as O

    threads = O.lib().oq_max_threads()
    O.lanczos_ground_cpu(O.ComplexCsr(O.Model(O.CHAIN, 10, 5, 1.0, 1.0, 0.2, 0.2)), nthreads=threads)  # warm-up
    sizes, counts, sample = [18, 20, 22], [64, 64, 128], [3, 2, 1]
    per, assembly = {}, {}
    k0 = 0
    for N, cnt, ns in zip(sizes, counts, sample):
        m = O.Model(O.CHAIN, N, N // 2, 1.0, 1.0, 0.2, 0.2)
        ts = []
        for j in range(ns):
            hi = O.rd_hi(1 + k0 + j, N, -2.0, 2.0)
            t = time.perf_counter()
            A = O.ComplexCsr(m, hi, 7)
            ta = time.perf_counter() - t
            t = time.perf_counter()
            O.lanczos_ground_cpu(A, tol=1e-9, nthreads=threads)
            ts.append(time.perf_counter() - t)
            assembly[N] = round(ta, 3)
            del A
        per[N] = float(np.mean(ts))
        k0 += cnt
    total = sum(per[N] * c for N, c in zip(sizes, counts))
    return {"value": round(len(sizes) and sum(counts) / total, 3), "unit": "realisations/s", "cores": threads,
            "kind": "port",
            "sample": f"{sum(sample)} realisations ({', '.join(f'{n} of N={N}' for n, N in zip(sample, sizes))}) "
                      "solved with the oracle's complex128 CSR (PETSc's layout) and a CPU three-term Lanczos to "
                      "1e-9, scaled to the 64/64/128 mix; assembly per lattice not included",
            "s_per_realisation": {str(N): round(per[N], 3) for N in sizes},
            "assembly_s": {str(N): assembly[N] for N in sizes}}


def chain16(reps: int = 5):
    """Periodic chain N=16, J1=D1=1, Sz=0 (D = 12,870): Krylov-Schur nev=1,
    tol 1e-9, E0 and <S^z_0 S^z_r> for all r; wall time of SolverInit +
    solve + solution + the correlators, after one untimed run."""
    import numpy as np
    import torch

    import lsquantumed_amd as Q

    env = Q.Environment(16)
    h = Q.Hamiltonian(env, Q.Basis(env, 0), Q.chain1D(env), 1.0, 1.0, 0.0, 0.0)
    h.build_diag(env)
    h.build_off_diag(env)
    psi = torch.empty(h.hamilt_size(), dtype=torch.float64, device="cuda")

    def run():
        eps = Q.Solver.SolverInit(h, 1, tol=1e-9)
        Q.Solver.solve(env, eps)
        lam, err = Q.Solver.solution(eps, 0, psi)
        Q.Solver.SolverClean(eps)
        return lam, err, Q.correlations(h, psi)["SzSz"][0]

    run()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        lam, err, zz = run()
    dt = (time.perf_counter() - t) / reps
    return {"workload": "periodic chain N=16, J1=D1=1, Sz=0, D=12870, Krylov-Schur nev=1 tol 1e-9",
            "E0": lam, "relative_error": err, "SzSz_0r": [float(v) for v in np.asarray(zz)],
            "wall_s": round(dt, 5)}


def hv_c128(h, B: int, real_ms: float, reps: int, warmup: int):
    """The complex128 product (lsqed_hv_c128: interleaved complex vectors, the
    layout of the complex-scalar PETSc Vec the reference runs on) on the same
    sector as the real one: H is real, both parts are updated in one pass, so the
    algorithmic bytes are twice the real product's.  HIP events on the launch
    stream, as for the real product; the real part of the result equals the
    real product of the real part bit for bit (checked here on the first
    launch)."""
    import torch

    D = h.hamilt_size()
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.complex(torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5,
                      torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5)
    y = torch.empty_like(x)
    h.matvec_c128(x, y)
    yr = h.matvec(x.real.contiguous())
    same = bool(torch.equal(y.real, yr))
    del yr
    stream = torch.cuda.current_stream()
    for _ in range(max(0, warmup - 1)):
        h.matvec_c128(x, y)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        h.matvec_c128(x, y)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    return {"workload": f"complex128 H.v (lsqed_hv_c128) on the same sector, D={D}, 16 B per entry",
            "hv_ms": round(ms, 3), "algorithmic_bytes": 2 * B, "effective_gbs": round(2 * B / ms / 1e6, 1),
            "time_vs_real_product": round(ms / real_ms, 3), "real_part_equals_real_product": same}


def hv_4x6_with_cpu():
    """H.v on the periodic 4x6 J1-J2 sector (D = 2,704,156): GPU (HIP events,
    the automatically chosen kernel) and CPU (the oracle's complex128 CSR in
    PETSc's layout, OpenMP SpMV over the host cores); GB/s by the algorithmic
    bytes B_Hv, and for the CPU also with the CSR bytes (24 B per stored
    nonzero + 16 B per complex gather and per row)."""
    import numpy as np
    import torch

    import lsquantumed_amd as Q
    from oracle import oracle as O

    env = Q.Environment(24)
    h = Q.Hamiltonian(env, Q.Basis(env, 0), Q.square2D(env, 4, 6), 1.0, 1.0, 0.5, 0.5)
    h.build_diag(env)
    h.build_off_diag(env)
    B, D, nnz = algorithmic_bytes(24, 12, 96)
    x = torch.rand(D, dtype=torch.float64, device="cuda") - 0.5
    y = torch.empty_like(x)
    h.matvec(x, y)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(20):
        h.matvec(x, y)
    ev1.record()
    torch.cuda.synchronize()
    gpu_ms = ev0.elapsed_time(ev1) / 20
    threads = O.lib().oq_max_threads()
    t = time.perf_counter()
    A = O.ComplexCsr(O.Model(O.SQUARE, 24, 12, 1.0, 1.0, 0.5, 0.5, lx=4, ly=6))
    build_s = time.perf_counter() - t
    xc = (np.random.default_rng(0).random(D) - 0.5).astype(np.complex128)
    yc = np.empty_like(xc)
    A.matvec(xc, yc, threads)
    t = time.perf_counter()
    for _ in range(3):
        A.matvec(xc, yc, threads)
    cpu_s = (time.perf_counter() - t) / 3
    csr_bytes = 24 * A.nnz + 16 * A.nnz + 16 * D
    return {"workload": "H.v, periodic square 4x6 (N=24), J1=D1=1, J2=D2=0.5, Sz=0, D=2704156",
            "algorithmic_bytes": B, "gpu_ms": round(gpu_ms, 4), "gpu_effective_gbs": round(B / gpu_ms / 1e6, 1),
            "gpu_kernel": {1: "generic", 2: "row split", 3: "row split (hipRTC)"}[h.hv_kernel()],
            "cpu_ms": round(cpu_s * 1e3, 2), "cpu_gbs": round(B / cpu_s / 1e9, 2),
            "cpu_csr_gbs": round(csr_bytes / cpu_s / 1e9, 2), "cpu_threads": threads,
            "cpu_csr_nnz": A.nnz, "cpu_csr_build_s": round(build_s, 2),
            "cpu_kind": "oracle complex128 CSR SpMV (PETSc's layout), OpenMP"}


def autograd_chain20(B: int = 64, N: int = 20):
    """Differentiable ground energy: 64 periodic chains of 20 spins, J1=1,
    J2=0.3 (D_k = J_k), h_i ~ U[-1, 1] from mt19937_64(1 + k); wall time of
    forward (batched solve with ground states) + backward (batched
    observables kernel), after one untimed pass; central differences
    (step 1e-5) of sum(E0) in J1 and J2 against the gradient."""
    import numpy as np
    import torch

    from lsquantumed_amd import ground_energy
    from lsquantumed_amd.tools import fields_from_seed

    h0 = np.stack([fields_from_seed(1 + k, N, -1.0, 1.0) for k in range(B)])

    def fwd_bwd():
        h = torch.tensor(h0, dtype=torch.float64, requires_grad=True)
        J1 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
        J2 = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
        E = ground_energy(h, J1, J2)
        E.sum().backward()
        return E, h.grad, J1.grad.item(), J2.grad.item()

    fwd_bwd()
    t = time.perf_counter()
    E, gh, gj1, gj2 = fwd_bwd()
    wall = time.perf_counter() - t
    d = 1e-5
    hh = torch.tensor(h0)
    with torch.no_grad():
        fj1 = (ground_energy(hh, 1.0 + d, 0.3).sum() - ground_energy(hh, 1.0 - d, 0.3).sum()).item() / (2 * d)
        fj2 = (ground_energy(hh, 1.0, 0.3 + d).sum() - ground_energy(hh, 1.0, 0.3 - d).sum()).item() / (2 * d)
    return {"workload": f"E0(h, J1, J2) and its gradient, {B} periodic chains of N={N}, J1=1, J2=0.3, "
                        "h_i~U[-1,1] from mt19937_64(1+k)",
            "forward_backward_s": round(wall, 4), "E0_mean": float(E.detach().mean()),
            "dE_dJ1": gj1, "dE_dJ1_fd": fj1, "dE_dJ1_rel_err": abs(gj1 - fj1) / abs(gj1),
            "dE_dJ2": gj2, "dE_dJ2_fd": fj2, "dE_dJ2_rel_err": abs(gj2 - fj2) / abs(gj2),
            "sum_dE_dh": float(gh.sum())}


def krylov_schur_4x6():
    """Krylov-Schur nev=10, ncv=32, tol 1e-9 on the periodic 4x6 square
    lattice, J1=D1=1, J2=D2=0.5, Sz=0 (D = 2,704,156): wall time of
    SolverInit + solve after one untimed solve, with SLEPc's stopping rule and
    with the multiplicity check (which finds both copies of the degenerate
    levels among the 10 lowest)."""
    import torch

    import lsquantumed_amd as Q

    env = Q.Environment(24)
    h = Q.Hamiltonian(env, Q.Basis(env, 0), Q.square2D(env, 4, 6), 1.0, 1.0, 0.5, 0.5)
    h.build_diag(env)
    h.build_off_diag(env)
    out = {"workload": "Krylov-Schur nev=10 ncv=32 tol 1e-9, periodic square 4x6 (N=24), J1=D1=1, J2=D2=0.5, "
                       "Sz=0, D=2704156"}
    for name, rounds in (("slepc_rule", 0), ("multiplicity_check", 3)):
        for rep in range(2):
            torch.cuda.synchronize()
            t = time.perf_counter()
            eps = Q.Solver.SolverInit(h, 10, 32, -2, tol=1e-9, multiplicity_rounds=rounds)
            nconv = Q.Solver.solve(env, eps)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            vals = [Q.Solver.solution(eps, n)[0] for n in range(min(nconv, 10))]
            its = eps.iterations()
            Q.Solver.SolverClean(eps)
        out[name] = {"solve_s": round(dt, 4), "restarts": its, "nconv": nconv, "eigenvalues": vals}
    return out


def krylov_schur_4x8(tol: float = 1e-9):
    """Krylov-Schur nev=4, ncv=20 on the periodic 4x8 square lattice,
    J1=D1=1, J2=D2=0.5, Sz=0 (D = 601,080,390): the basis is 21 vectors of
    4.81 GB.  Reports the eigenvalues, EPS_ERROR_RELATIVE of each, the
    restarts and the wall time, and the ground energy of the three-vector
    Lanczos on the same problem as a cross-check."""
    import torch

    import lsquantumed_amd as Q

    env = Q.Environment(32)
    h = Q.Hamiltonian(env, Q.Basis(env, 0), Q.square2D(env, 4, 8), 1.0, 1.0, 0.5, 0.5)
    h.build_diag(env)
    h.build_off_diag(env)
    log("4x8 Krylov-Schur nev=4 ncv=20: start")
    torch.cuda.synchronize()
    t = time.perf_counter()
    eps = Q.Solver.SolverInit(h, 4, 20, -2, tol=tol)
    nconv = Q.Solver.solve(env, eps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    sol = [Q.Solver.solution(eps, n) for n in range(min(nconv, 4))]
    its = eps.iterations()
    Q.Solver.SolverClean(eps)
    log(f"4x8 Krylov-Schur: {dt:.1f} s, {its} restarts, nconv {nconv}")
    del eps
    torch.cuda.empty_cache()
    # the solve may have run on a renumbered copy of the sites: prepare the
    # model's own kernel before the timed Lanczos, which uses it
    h.hv_kernel()
    r = Q.solver.lanczos_ground_residual(h, tol=1e-11, maxit=800, seed=5)
    e0, steps, dl = r["E0"], r["steps"], r["wall_s"]
    torch.cuda.empty_cache()
    return {"workload": f"Krylov-Schur nev=4 ncv=20 tol {tol:g}, periodic square 4x8 (N=32), J1=D1=1, J2=D2=0.5, "
                        "Sz=0, D=601080390 (basis 21 x 4.81 GB)",
            "solve_s": round(dt, 2), "restarts": its, "nconv": nconv,
            "eigenvalues": [v for v, _ in sol], "relative_errors": [e for _, e in sol],
            "lanczos3_E0": e0, "lanczos3_steps": steps, "lanczos3_s": round(dl, 2),
            "lanczos3": "three-vector Lanczos to a 1e-11 relative residual, the cross-check of the lowest value",
            "E0_diff_krylov_schur_vs_lanczos3": abs(sol[0][0] - e0) if sol else None}


def lanczos_6x6(maxit: int = 600, tol: float = 1e-9):
    """Ground state of the periodic 6x6 square lattice, J1=D1=1, J2=D2=0.5,
    Sz=0 (N=36, D = 9,075,135,300, 72.6 GB per vector) with the three-vector
    Lanczos: q, p and y on the GPU (218 GB), one lsqed_lanczos_step per step
    (H.v with the alpha dot fused, then the fused update and norm), the
    tridiagonal matrix on the host; stops when the lowest Ritz pair has
    |beta_m s_m| / |theta| < tol.  Progress lines go to stderr every 5 steps.
    Also times one H.v alone (HIP events) and reports its HBM traffic when the
    counter record of this kernel exists."""
    import torch

    import lsquantumed_amd as Q

    N = 36
    env = Q.Environment(N)
    h = Q.Hamiltonian(env, Q.Basis(env, 0), Q.square2D(env, 6, 6), 1.0, 1.0, 0.5, 0.5)
    h.build_diag(env)
    h.build_off_diag(env)
    D = h.hamilt_size()
    kernel = {1: "generic", 2: "row split", 3: "row split (hipRTC)"}[h.hv_kernel()]
    fp = h.hv_fingerprint()
    x = torch.empty(D, dtype=torch.float64, device="cuda").uniform_(-0.5, 0.5)
    y = torch.empty(D, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    h.matvec(x, y)  # warm-up (kernel compilation, tables)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    h.matvec(x, y)
    ev1.record(stream)
    torch.cuda.synchronize()
    hv_ms = ev0.elapsed_time(ev1)
    del x, y
    torch.cuda.empty_cache()
    B, _, nnz = algorithmic_bytes(N, N // 2, 144)
    rec = measured_traffic(fp)
    log(f"6x6: D={D} kernel={kernel} H.v {hv_ms:.1f} ms, {B / hv_ms / 1e6:.1f} GB/s effective")
    r = Q.solver.lanczos_ground_residual(
        h, tol=tol, maxit=maxit, seed=7,
        progress=lambda m, th, res, sec: log(f"6x6 step {m}: theta {th:.12f} resid {res:.3e} ({sec:.1f} s)"))
    steps = r["steps"]
    return {"workload": f"ground state, periodic square 6x6 (N=36), J1=D1=1, J2=D2=0.5, Sz=0, D={D}, "
                        f"three-vector Lanczos to a {tol:g} relative residual",
            "E0": r["E0"], "E0_per_site": r["E0"] / N, "steps": steps, "relative_residual": r["relative_residual"],
            "wall_s": round(r["wall_s"], 2), "s_per_step": round(r["wall_s"] / steps, 4),
            "lanczos_steps_per_s": round(steps / r["wall_s"], 3), "hv_kernel": kernel, "hv_fingerprint": fp,
            "hv_ms": round(hv_ms, 2), "hv_algorithmic_bytes": B, "hv_effective_gbs": round(B / hv_ms / 1e6, 1),
            "hv_hbm_bytes": rec["hbm_bytes"] if rec else None,
            "hv_hbm_gbs": round(rec["hbm_bytes"] / hv_ms / 1e6, 1) if rec else None,
            "hv_roofline_frac": round(rec["hbm_bytes"] / hv_ms / 1e6 / HBM_PEAK_GBS, 4) if rec else None,
            "device_gb": round(3 * D * 8 / 1e9, 1)}


def hv_only(lx: int, ly: int, reps: int, warmup: int, chain: int = 0, hv_kernel: int = 0):
    """H.v alone with two device vectors, timed with HIP events over `reps`
    launches on the launch stream; x.(Hx) is reported as a checksum.  The
    periodic lx x ly square lattice (J1=D1=1, J2=D2=0.5), or with `chain` > 0
    the periodic chain of that many spins (J1=D1=1, J2=D2=0.2, the disorder
    workload's couplings, no fields); Sz=0.  The row split and block size
    follow LSQED_SPLIT_L / LSQED_JIT_ROW_THREADS when set."""
    import torch

    import lsquantumed_amd as Q

    N = chain if chain > 0 else lx * ly
    env = Q.Environment(N)
    if chain > 0:
        lat = Q.chain1D(env)
        h = Q.Hamiltonian(env, Q.Basis(env, 0), lat, 1.0, 1.0, 0.2, 0.2)
    else:
        lat = Q.square2D(env, lx, ly)
        h = Q.Hamiltonian(env, Q.Basis(env, 0), lat, 1.0, 1.0, 0.5, 0.5)
    h.build_diag(env)
    h.build_off_diag(env)
    if hv_kernel:
        h.set_hv_kernel(hv_kernel)
    D = h.hamilt_size()
    B, _, _ = algorithmic_bytes(N, N // 2, len(lat.nn_bonds[0]) + len(lat.nnn_bonds[0]))
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(D, dtype=torch.float64, device="cuda", generator=g)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    for _ in range(max(1, warmup)):
        h.matvec(x, y)
    torch.cuda.synchronize()
    chk = sum(float(torch.dot(x[i:i + 2**30], y[i:i + 2**30]).item()) for i in range(0, D, 2**30))
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        h.matvec(x, y)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    kernel = {1: "generic", 2: "row split", 3: "row split (hipRTC)"}[h.hv_kernel()]
    fp = h.hv_fingerprint()
    rec = measured_traffic(fp)
    name = f"periodic chain N={N}, J1=D1=1, J2=D2=0.2" if chain > 0 else \
        f"periodic square {lx}x{ly} (N={N}), J1=D1=1, J2=D2=0.5"
    return {"workload": f"H.v, {name}, Sz=0, D={D}",
            "hv_kernel": kernel, "fingerprint": fp, "split_low_bits": os.environ.get("LSQED_SPLIT_L", "default"),
            "row_threads": os.environ.get("LSQED_JIT_ROW_THREADS", "default"), "hv_ms": round(ms, 4),
            "algorithmic_bytes": B, "effective_gbs": round(B / ms / 1e6, 1),
            "hbm_bytes": rec["hbm_bytes"] if rec else None,
            "hbm_gbs": round(rec["hbm_bytes"] / ms / 1e6, 1) if rec else None,
            "roofline_frac": round(rec["hbm_bytes"] / ms / 1e6 / HBM_PEAK_GBS, 4) if rec else None,
            "x_dot_hx": chk}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lx", type=int, default=4)
    ap.add_argument("--ly", type=int, default=8)
    ap.add_argument("--hv-kernel", type=int, default=0, help="0 auto, 1 generic, 2 row split, 3 specialised row split")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-disorder measurement")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched-disorder measurement (one JSON object; used for profiling)")
    ap.add_argument("--lanczos-6x6", action="store_true",
                    help="only the 6x6 (N=36) three-vector Lanczos ground state (218 GB; minutes)")
    ap.add_argument("--maxit", type=int, default=600,
                    help="with --lanczos-6x6: at most this many Lanczos steps (a few for a profile)")
    ap.add_argument("--krylov-4x8", action="store_true",
                    help="only Krylov-Schur nev=4 ncv=20 on the 4x8 sector (101 GB basis + a 4.8 GB spare; about a minute)")
    ap.add_argument("--hv-only", action="store_true",
                    help="only H.v on the --lx x --ly lattice with two vectors (6x6: 145 GB); one JSON object")
    ap.add_argument("--chain", type=int, default=0, help="with --hv-only: a periodic chain of this many spins")
    ap.add_argument("--stream-sizes", action="store_true",
                    help="only the read-stream bandwidth against the streamed buffer's size (16 MB - 4 GB)")
    ap.add_argument("--calibrate", action="store_true",
                    help="only the kernels of known traffic used to read the HBM byte counters (for rocprofv3 --pmc)")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 4x8 Krylov-Schur (nev=4, ncv=20) and the 6x6 Lanczos legs of the default run")
    ap.add_argument("--record-traffic", metavar="PREFIX",
                    help="store the H.v traffic record built from rocprofv3 --pmc passes PREFIX_<i>/ (made with "
                         "--hv-only) in profiles/hv_traffic.json, keyed by the kernel fingerprint; no GPU use")
    args = ap.parse_args()
    if args.record_traffic:
        fp, rec = record_traffic(args.record_traffic)
        print(json.dumps({"fingerprint": fp, **rec}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the Sz-sector path here is built and measured on a single "
              "MI355X (multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)

    import torch

    import lsquantumed_amd as Q

    torch.cuda.set_device(0)
    if args.calibrate:
        print(json.dumps({"calibrate": calibrate()}))
        return
    if args.stream_sizes:
        print(json.dumps({"stream_sizes": stream_sizes()}))
        return
    if args.batch_only:
        print(json.dumps({"batched_disorder": disorder_batch()}))
        return
    if args.lanczos_6x6:
        print(json.dumps({"lanczos_6x6": lanczos_6x6(maxit=max(1, args.maxit))}))
        return
    if args.krylov_4x8:
        print(json.dumps({"krylov_schur_4x8": krylov_schur_4x8()}))
        return
    if args.hv_only:
        print(json.dumps({"hv": hv_only(args.lx, args.ly, args.steps, args.warmup, args.chain, args.hv_kernel)}))
        return
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
    fingerprint = h.hv_fingerprint()
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
    rec = measured_traffic(fingerprint)
    traffic = rec["hbm_bytes"] if rec else None

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

    stream = stream_ceiling(q, p, y)
    del q, p, y
    torch.cuda.empty_cache()
    c128 = hv_c128(h, B, hv_ms, args.steps, args.warmup)
    torch.cuda.empty_cache()
    batch = None if args.no_batch else disorder_batch()
    if batch is not None and not args.no_cpu:
        batch["cpu_baseline"] = cpu_disorder_baseline()
    ks = None if args.no_batch else krylov_schur_4x6()
    c1 = None if args.no_batch else chain16()
    ag = None if args.no_batch else autograd_chain20()
    hv46 = None if (args.no_batch or args.no_cpu) else hv_4x6_with_cpu()
    del h
    torch.cuda.empty_cache()
    ks48 = None if args.no_large else krylov_schur_4x8()
    torch.cuda.empty_cache()
    l66 = None if args.no_large else lanczos_6x6()
    torch.cuda.empty_cache()
    cpu = None if args.no_cpu else cpu_baseline((N, N // 2, J1, D1, J2, D2, lx, ly, B))
    hbm_gbs = traffic / (hv_ms * 1e-3) / 1e9 if traffic else None
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
        # achieved / frac: the HBM bytes one launch of this exact kernel moves (hardware counters of
        # the committed record whose fingerprint matches, profiles/hv_traffic.json) over the launch
        # time measured here; null when no record of this kernel exists.  The algorithmic figure
        # counts one 8-byte read per gather, also those served by a cache, so it can exceed the
        # peak: it is `value`, the effective bandwidth, not a roofline fraction.
        "roofline": {"bound": "hbm", "achieved": round(hbm_gbs, 1) if hbm_gbs else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(hbm_gbs / HBM_PEAK_GBS, 4) if hbm_gbs else None,
                     "traffic": traffic, "kernel": "hv_jit_nodot" if kernel_name == "row split (hipRTC)" else kernel_name,
                     "fingerprint": fingerprint,
                     "traffic_read_bytes": rec["read_bytes"] if rec else None,
                     "traffic_write_bytes": rec["write_bytes"] if rec else None,
                     "algorithmic_bytes": B, "effective_gbs": round(hv_gbs, 1),
                     "frac_of_measured_read_stream": round(hbm_gbs / stream["read_gbs"], 4) if hbm_gbs else None,
                     "measured_read_stream_gbs": stream["read_gbs"],
                     "measured_triad_gbs": max(stream["triad_gbs"], stream["torch_triad_gbs"])},
        "stream": stream,
        "cpu_baseline": cpu,
        "batched_disorder": batch,
        "krylov_schur_4x6": ks,
        "hv_c128": c128,
        "hv_4x6": hv46,
        "krylov_schur_4x8": ks48,
        "lanczos_6x6": l66,
        "chain16": c1,
        "autograd_chain20": ag,
        "alpha_last": alphas[-1],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
