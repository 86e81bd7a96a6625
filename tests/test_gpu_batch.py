"""GPU parity of the batched disorder solver and the differentiable E0.

Batched ground energies against the oracle's dense spectra (small chains)
and against the plain one-at-a-time Krylov-Schur solve (the loop of
examples/main_disorder.cpp:82-120) on samples of the N = 18/20/22 disorder
workload; ground states and their Hellmann-Feynman observables against the
oracle's eigenvectors and correlators; autograd gradients against central
finite differences (step 1e-5) at 1e-6 relative."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLAGS_DIS = 1 | 2 | 4  # build_diag + build_off_diag + build_disorder


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available()
    return T


def oracle_ground(O, N, J1, J2, hi, vectors=False):
    m = O.Model(O.CHAIN, N, N // 2, J1, J1, J2, J2)
    A = m.csr(hi, FLAGS_DIS)
    w, v = O.eigs(A, 1)
    return (w[0], v[:, 0]) if vectors else w[0]


def plain_e0(h, tol=1e-12):
    from lsquantumed_amd import Solver
    eps = Solver.SolverInit(h, 1, tol=tol)
    nconv = Solver.solve(None, eps)
    assert nconv >= 1
    lam, err = Solver.solution(eps, 0)
    Solver.SolverClean(eps)
    return lam


def test_batch_small_chains_vs_oracle(O, torch):
    """Mixed N = 8/10/12 chains with J2 and random fields: every E0 equals the
    lowest eigenvalue of the oracle's matrix (dense) to 1e-11 relative."""
    from lsquantumed_amd.batch import BatchSolver, disorder_chains
    sizes, counts = [8, 10, 12], [3, 3, 4]
    hams = disorder_chains(sizes, counts, J1=1.0, J2=0.2, W=2.0, seed=1)
    say("batch small", sizes, counts)
    b = BatchSolver(hams, tol=1e-12)
    e0 = b.solve()
    for k, h in enumerate(hams):
        ref = oracle_ground(O, h.basis.nspins, 1.0, 0.2, h.hi)
        assert abs(e0[k] - ref) <= 1e-11 * abs(ref), (k, e0[k], ref)
        steps, res, conv = b.info(k)
        assert conv and res < 1e-12 and steps >= 2
    sched = b.schedule()
    assert sched["lanes_per_state"] in (1, 2, 4) and sched["rows_per_task"] > 0


def test_batch_disorder_sample_vs_plain(O, torch):
    """Sample of the N = 18/20/22 workload (J1 = 1, J2 = 0.2, h ~ U[-2, 2] from
    mt19937_64(1 + k)): batched E0 = plain Krylov-Schur E0 to 1e-9 relative,
    and one N = 18 value against the oracle's sparse eigensolve."""
    from lsquantumed_amd.batch import BatchSolver, disorder_chains
    sizes, counts = [18, 20, 22], [4, 4, 8]
    hams = disorder_chains(sizes, counts, J1=1.0, J2=0.2, W=2.0, seed=1)
    say("batch sample", sizes, counts)
    b = BatchSolver(hams, tol=1e-9)
    e0 = b.solve()
    for k, h in enumerate(hams):
        ref = plain_e0(h)
        assert abs(e0[k] - ref) <= 1e-9 * abs(ref), (k, e0[k], ref)
        assert b.info(k)[2]
    ref0 = oracle_ground(O, 18, 1.0, 0.2, hams[0].hi)
    assert abs(e0[0] - ref0) <= 1e-9 * abs(ref0)


@pytest.mark.timeout(600)
def test_batch_full_disorder_workload(O, torch):
    """The whole 256-realisation workload (64 chains of N = 18, 64 of 20, 128
    of 22; J1 = D1 = 1, J2 = D2 = 0.2, h ~ U[-2, 2] from mt19937_64(1 + k)),
    solved as one batch to a 1e-9 residual: every E0 equals the plain
    one-at-a-time Krylov-Schur solve (the loop of examples/main_disorder.cpp:
    82-120) to 1e-9 relative, and the first and last realisation of each N
    equal the oracle's sparse eigensolve of the assembled matrix to 1e-9."""
    from lsquantumed_amd import Solver
    from lsquantumed_amd.batch import BatchSolver, disorder_chains
    sizes, counts = [18, 20, 22], [64, 64, 128]
    hams = disorder_chains(sizes, counts, J1=1.0, J2=0.2, W=2.0, seed=1)
    say("full batch", sizes, counts)
    b = BatchSolver(hams, tol=1e-9)
    e0 = b.solve()
    assert all(b.info(k)[2] for k in range(len(hams)))
    say("plain loop")
    worst = 0.0
    for k, h in enumerate(hams):
        eps = Solver.SolverInit(h, 1, tol=1e-9)
        assert Solver.solve(None, eps) >= 1, k
        lam, _ = Solver.solution(eps, 0)
        Solver.SolverClean(eps)
        worst = max(worst, abs(e0[k] - lam) / abs(lam))
    say("max rel diff vs plain", worst)
    assert worst <= 1e-9
    first = 0
    for N, c in zip(sizes, counts):
        for k in (first, first + c - 1):
            assert hams[k].basis.nspins == N
            say("oracle", N, k)
            ref = oracle_ground(O, N, 1.0, 0.2, hams[k].hi)
            assert abs(e0[k] - ref) <= 1e-9 * abs(ref), (N, k, e0[k], ref)
        first += c
    b.destroy()


def test_batch_schedules_agree(torch):
    """Every H.v schedule gives the same E0 (the Lanczos recurrence is the
    same up to rounding): the chunk-table kernel at 1, 2 and 4 lanes per
    state, the row split specialised per lattice group, and the autotuned
    choice; N = 18 and 20 exercise the specialised kernel's high bits."""
    from lsquantumed_amd.batch import BatchSolver, disorder_chains
    hams = disorder_chains([12, 18, 20], [3, 2, 2], J1=1.0, J2=0.2, W=2.0, seed=7)
    vals = []
    for kern, lps in ((0, 0), (1, 1), (1, 2), (1, 4), (2, 0)):
        say("schedule", kern, lps)
        b = BatchSolver(hams, tol=1e-12, lanes_per_state=lps, hv_kernel=kern)
        vals.append(b.solve())
        sch = b.schedule()
        say("  ", sch)
        if lps:
            assert sch["lanes_per_state"] == lps and sch["hv_kernel"] == "chunk table"
        if kern == 2:
            assert sch["hv_kernel"] == "specialised row split"
        b.destroy()
    for v in vals[1:]:
        np.testing.assert_allclose(v, vals[0], rtol=1e-12, atol=0)


def test_batch_group_chains_agree_with_lockstep(O, torch, monkeypatch):
    """One Lanczos chain per lattice group on its own stream (the default when
    each group's problems are contiguous) gives bit for bit the energies and
    step counts of the lockstep replay (LSQED_BATCH_GROUP_CHAINS=0), with and
    without ground-state vectors; a batch whose groups interleave runs the
    lockstep replay and still meets the oracle."""
    from lsquantumed_amd.batch import BatchSolver, disorder_chains
    hams = disorder_chains([12, 14, 16], [3, 2, 3], J1=1.0, J2=0.2, W=2.0, seed=11)
    runs = {}
    for mode in ("chains", "lockstep"):
        if mode == "lockstep":
            monkeypatch.setenv("LSQED_BATCH_GROUP_CHAINS", "0")
        for kv in (False, True):
            say("group chains", mode, "keep_vectors", kv)
            b = BatchSolver(hams, tol=1e-12, hv_kernel=BatchSolver.HV_SPECIALISED, keep_vectors=kv)
            e0 = b.solve()
            runs[(mode, kv)] = (e0.copy(), [b.info(k)[0] for k in range(len(hams))])
            b.destroy()
    monkeypatch.delenv("LSQED_BATCH_GROUP_CHAINS", raising=False)
    for kv in (False, True):
        assert np.array_equal(runs[("chains", kv)][0], runs[("lockstep", kv)][0]), kv
        assert runs[("chains", kv)][1] == runs[("lockstep", kv)][1], kv
    for k, h in enumerate(hams):
        ref = oracle_ground(O, h.basis.nspins, 1.0, 0.2, h.hi)
        assert abs(runs[("chains", False)][0][k] - ref) <= 1e-11 * abs(ref), (k, ref)
    mixed = [hams[0], hams[3], hams[1], hams[5], hams[4]]  # N = 12, 14, 12, 16, 14: groups interleave
    b = BatchSolver(mixed, tol=1e-12, hv_kernel=BatchSolver.HV_SPECIALISED)
    e0 = b.solve()
    b.destroy()
    for k, h in enumerate(mixed):
        ref = oracle_ground(O, h.basis.nspins, 1.0, 0.2, h.hi)
        assert abs(e0[k] - ref) <= 1e-11 * abs(ref), (k, ref)


def test_batch_low_bond_record_modes_agree(O, torch, monkeypatch):
    """The batched H.v's low bonds by rank arithmetic (LSQED_JIT_LOWTAB=0),
    from full per-pattern records (1) or from records of the bonds needing
    two rank lookups only (2, the default): every bond's term is summed in
    bond order either way, so the energies and step counts agree bit for bit."""
    from lsquantumed_amd.batch import BatchSolver, disorder_chains
    hams = disorder_chains([14, 18, 20], [2, 2, 2], J1=1.0, J2=0.2, W=2.0, seed=5)
    runs = []
    for mode in ("0", "1", "2"):
        monkeypatch.setenv("LSQED_JIT_LOWTAB", mode)
        say("low-bond records mode", mode)
        b = BatchSolver(hams, tol=1e-12, hv_kernel=BatchSolver.HV_SPECIALISED)
        runs.append((b.solve().copy(), [b.info(k)[0] for k in range(len(hams))]))
        b.destroy()
    monkeypatch.delenv("LSQED_JIT_LOWTAB", raising=False)
    for e0, steps in runs[1:]:
        assert np.array_equal(e0, runs[0][0]) and steps == runs[0][1]
    for k, h in enumerate(hams):
        ref = oracle_ground(O, h.basis.nspins, 1.0, 0.2, h.hi)
        assert abs(runs[0][0][k] - ref) <= 1e-11 * abs(ref), (k, ref)


def test_batch_resolve_after_new_fields(O, torch):
    """Fields set on the models between solves take effect at the next solve
    (the reference rebuilds the disorder per realisation, main_disorder.cpp:83-84)."""
    from lsquantumed_amd import fields_from_seed
    from lsquantumed_amd.batch import BatchSolver, disorder_chains
    hams = disorder_chains([10, 12], [2, 2], J1=1.0, J2=0.2, W=2.0, seed=3)
    b = BatchSolver(hams, tol=1e-12)
    b.solve()
    for k, h in enumerate(hams):
        h.build_disorder(None, fields_from_seed(100 + k, h.basis.nspins, -1.0, 1.0))
    e0 = b.solve()
    for k, h in enumerate(hams):
        ref = oracle_ground(O, h.basis.nspins, 1.0, 0.2, h.hi)
        assert abs(e0[k] - ref) <= 1e-11 * abs(ref)


def test_batch_vectors_and_observables(O, torch):
    """Ground states of the second pass: unit norm, eigenvectors of H
    (||H psi - E0 psi|| <= 1e-9), equal to the oracle's up to sign; the
    observables kernel gives <S^z_i> and the bond sums of <S^z S^z> and
    <(S^+S^- + S^-S^+)/2> of the oracle's correlators (SzSz/SpSm::ExpectVal)
    to 1e-9, and E0 = J1 dE/dJ1 + J2 dE/dJ2 + sum_i h_i dE/dh_i (H is linear
    in its parameters)."""
    from lsquantumed_amd.batch import (OBS_FLIP_NN, OBS_FLIP_NNN, OBS_NORM, OBS_ZZ_NN, OBS_ZZ_NNN, BatchSolver,
                                       disorder_chains)
    J1, J2 = 1.0, 0.3
    hams = disorder_chains([10, 12], [2, 3], J1=J1, J2=J2, W=1.0, seed=11)
    b = BatchSolver(hams, tol=1e-12, keep_vectors=True)
    e0 = b.solve()
    say("schedule", b.schedule())
    obs = b.observables()
    for k, h in enumerate(hams):
        N = h.basis.nspins
        nup = N // 2
        say("vectors", k, N)
        psi = b.ground_state(k)
        assert abs(torch.linalg.vector_norm(psi).item() - 1.0) < 1e-12
        r = h.matvec(psi) - e0[k] * psi
        assert torch.linalg.vector_norm(r).item() <= 1e-9 * abs(e0[k])
        ref_e, ref_v = oracle_ground(O, N, J1, J2, h.hi, vectors=True)
        p = psi.cpu().numpy()
        assert abs(abs(p @ ref_v) - 1.0) < 1e-10
        # <S^z_i>
        states = O.int_basis(N, nup)
        w = ref_v ** 2
        sz = np.array([np.sum(w * np.where((states >> i) & 1, 0.5, -0.5)) for i in range(N)])
        np.testing.assert_allclose(obs[k, :N], sz, rtol=0, atol=1e-9)
        assert abs(obs[k, OBS_NORM] - 1.0) < 1e-12
        # bond sums from the oracle's correlators
        for kind, (zz_i, fl_i) in enumerate([(OBS_ZZ_NN, OBS_FLIP_NN), (OBS_ZZ_NNN, OBS_FLIP_NNN)]):
            bi, bj = O.bonds(O.CHAIN, N, kind=kind)
            zz = sum(O.szsz(N, nup, ref_v, int(i), int(j)) for i, j in zip(bi, bj))
            fl = sum(0.5 * (O.spsm(N, nup, ref_v, int(i), int(j)) + O.smsp(N, nup, ref_v, int(i), int(j)))
                     for i, j in zip(bi, bj))
            assert abs(obs[k, zz_i] - zz) < 1e-9, (k, kind, obs[k, zz_i], zz)
            assert abs(obs[k, fl_i] - fl) < 1e-9, (k, kind, obs[k, fl_i], fl)
        euler = (J1 * (obs[k, OBS_ZZ_NN] + obs[k, OBS_FLIP_NN]) + J2 * (obs[k, OBS_ZZ_NNN] + obs[k, OBS_FLIP_NNN])
                 + float(np.dot(h.hi, obs[k, :N])))
        assert abs(euler - e0[k]) <= 1e-10 * abs(e0[k])


def test_batch_errors(torch):
    import lsquantumed_amd as Q
    from lsquantumed_amd.batch import BatchSolver, disorder_chains
    hams = disorder_chains([8], [2], J1=1.0, J2=0.0, W=1.0)
    with pytest.raises(Q.LsqedError):
        BatchSolver(hams, lanes_per_state=3)
    with pytest.raises(ValueError):
        BatchSolver([])
    b = BatchSolver(hams)
    with pytest.raises(Q.LsqedError):
        b.info(0)  # not solved
    b.solve()
    with pytest.raises(Q.LsqedError):
        b.ground_state(0)  # no vectors kept
    with pytest.raises(Q.LsqedError):
        b.observables()
    with pytest.raises(Q.LsqedError):
        b.info(5)
    env = Q.Environment(34)
    big = Q.Hamiltonian(env, Q.Basis(env, 0), Q.chain1D(env), 1.0, 1.0, 0.0, 0.0)
    with pytest.raises(Q.LsqedError):
        BatchSolver([big])


def test_batch_workspace_bytes(torch):
    """lsqed_batch_workspace_bytes: 3 doubles per basis state of every model
    (4 with the ground-state vectors kept), -1 for an empty batch."""
    import ctypes as C

    from lsquantumed_amd import lib
    from lsquantumed_amd._lib import BatchOpts
    from lsquantumed_amd.batch import BatchSolver, disorder_chains
    hams = disorder_chains([8, 10, 12], [2, 1, 3], J1=1.0, J2=0.2, W=1.0)
    states = 2 * 70 + 252 + 3 * 924
    assert sum(h.hamilt_size() for h in hams) == states
    assert BatchSolver.workspace_bytes(hams) == 3 * 8 * states
    assert BatchSolver.workspace_bytes(hams, keep_vectors=True) == 4 * 8 * states
    assert lib().lsqed_batch_workspace_bytes(0, None, C.byref(BatchOpts(0, 0.0, 0, 0, 0, 0))) == -1


def _fd_check(torch, B, N, J1v, J2v, seed, fd_rows):
    from lsquantumed_amd.autograd import ground_energy
    from lsquantumed_amd.tools import fields_from_seed
    h0 = np.stack([fields_from_seed(seed + k, N, -1.0, 1.0) for k in range(B)])
    h = torch.tensor(h0, dtype=torch.float64, requires_grad=True)
    J1 = torch.tensor(J1v, dtype=torch.float64, requires_grad=True)
    J2 = torch.tensor(J2v, dtype=torch.float64, requires_grad=True)
    say("autograd forward", B, N)
    E = ground_energy(h, J1, J2)
    E.sum().backward()
    gh, gj1, gj2 = h.grad.numpy(), J1.grad.item(), J2.grad.item()
    d = 1e-5
    # dE_b/dh_bi for the chosen realisations: one batched solve of all +/- perturbations
    rows = []
    for b in fd_rows:
        for i in range(N):
            for s in (+1, -1):
                r = h0[b].copy()
                r[i] += s * d
                rows.append(r)
    say("fd fields", len(rows))
    with torch.no_grad():
        Ep = ground_energy(torch.tensor(np.array(rows)), J1v, J2v).numpy().reshape(len(fd_rows), N, 2)
    fd_h = (Ep[:, :, 0] - Ep[:, :, 1]) / (2 * d)
    for n, b in enumerate(fd_rows):
        err = np.linalg.norm(fd_h[n] - gh[b]) / np.linalg.norm(gh[b])
        say("dE/dh", b, err)
        assert err <= 1e-6, (b, err)
    with torch.no_grad():
        hh = torch.tensor(h0)
        fj1 = (ground_energy(hh, J1v + d, J2v).sum() - ground_energy(hh, J1v - d, J2v).sum()).item() / (2 * d)
        fj2 = (ground_energy(hh, J1v, J2v + d).sum() - ground_energy(hh, J1v, J2v - d).sum()).item() / (2 * d)
    say("dE/dJ", gj1, fj1, gj2, fj2)
    assert abs(fj1 - gj1) <= 1e-6 * abs(gj1)
    assert abs(fj2 - gj2) <= 1e-6 * abs(gj2)
    return E.detach().numpy(), h0


def test_autograd_small_fd(torch):
    """E0(h, J1, J2) gradients on 4 chains of 12 spins against central
    differences (step 1e-5): norm-wise relative error <= 1e-6."""
    _fd_check(torch, B=4, N=12, J1v=1.0, J2v=0.3, seed=5, fd_rows=[0, 1, 2, 3])


def test_autograd_chain20_batch64(O, torch):
    """The differentiable-E0 workload: 64 chains of 20 spins, J1 = 1,
    J2 = 0.3, h ~ U[-1, 1] (mt19937_64(1 + k)); gradients of the batch sum
    against central differences at 1e-6 relative (dE/dh for two
    realisations, dE/dJ1 and dE/dJ2 over the whole batch); E0 of two
    realisations against the plain Krylov-Schur solve at 1e-9."""
    from lsquantumed_amd.batch import disorder_chains
    E, h0 = _fd_check(torch, B=64, N=20, J1v=1.0, J2v=0.3, seed=1, fd_rows=[0, 37])
    hams = disorder_chains([20], [2], J1=1.0, J2=0.3, W=1.0, seed=1)
    for k in range(2):
        np.testing.assert_allclose(hams[k].hi, h0[k], rtol=0, atol=0)
        ref = plain_e0(hams[k])
        assert abs(E[k] - ref) <= 1e-9 * abs(ref)


def test_disorder_average_driver_and_resume(O, torch, tmp_path):
    """The realisation loop of main_disorder.cpp with batched ground states:
    every dsf_d<k>.dat has the ground energy of the oracle's dense spectrum,
    excitation energies of the oracle, and multiplet-summed DSF weights of
    the oracle's eigenvectors at every chain momentum; a deleted file is
    recomputed on resume and the others are left alone."""
    import os

    import lsquantumed_amd as Q
    from lsquantumed_amd.disorder import disorder_average
    from test_oracle import _parse_dsf
    N, nev = 8, 6
    env = Q.Environment(N)
    h = Q.Hamiltonian(env, Q.Basis(env, 0), Q.chain1D(env), 1.0, 1.0, 0.2, 0.2)
    h.build_diag(env)
    h.build_off_diag(env)
    rd = Q.RandomDisorder(N, True, -1.0, 1.0, 3)
    done = disorder_average(env, h, rd, str(tmp_path), nev=nev)
    assert sorted(done) == [0, 1, 2]
    coords = np.array([[i, 0] for i in range(N)], float)
    for k in range(3):
        hi = rd.hi_init(k).copy()
        m = O.Model(O.CHAIN, N, N // 2, 1.0, 1.0, 0.2, 0.2)
        w, v = np.linalg.eigh(m.csr(hi, FLAGS_DIS).toarray())
        gs, blocks = _parse_dsf(open(tmp_path / f"dsf_d{k}.dat").read())
        assert abs(gs - w[0]) < 1e-11 and abs(done[k] - w[0]) < 1e-11
        ex = np.array([e for e, _ in blocks[0]])
        np.testing.assert_allclose(ex, w[:nev] - w[0], rtol=0, atol=1e-10)
        amp = np.array([[blocks[i][n][1] for i in range(N)] for n in range(nev)])
        ref_amp = O.dsf_amplitudes(N, N // 2, v[:, 0], v[:, :nev], "z")
        groups = O.multiplets(w[:nev])
        if abs(w[nev] - w[nev - 1]) <= 1e-8:
            groups = groups[:-1]
        for grp in groups:
            for qx in range(N):
                q = (2 * np.pi * qx / N, 0.0)
                W = sum(O.dsf_weight(amp[n], q, coords) for n in grp)
                Wr = sum(O.dsf_weight(ref_amp[n], q, coords) for n in grp)
                assert abs(W - Wr) < 1e-9, (k, grp, qx, W, Wr)
    t0 = os.path.getmtime(tmp_path / "dsf_d0.dat")
    os.remove(tmp_path / "dsf_d1.dat")
    again = disorder_average(env, h, rd, str(tmp_path), nev=nev)
    assert sorted(again) == [1] and abs(again[1] - done[1]) < 1e-12
    assert os.path.getmtime(tmp_path / "dsf_d0.dat") == t0


def test_batch_mixed_lattices_and_sectors(O, torch):
    """One batch holding a 3x4 square lattice (Sz = 0 and Sz = 1), a 6x2
    honeycomb with J2, a 10-site chain in Sz = -1, an 8-site chain and a 2x4
    square lattice (doubled bonds), with and without fields: each model its
    own specialised-kernel group (six groups, more than the three side
    streams the smaller groups' launches share); every E0 equals the
    oracle's dense spectrum to 1e-11, for both H.v kernels."""
    import lsquantumed_amd as Q
    from lsquantumed_amd.batch import BatchSolver
    from test_gpu_hamiltonian import make
    specs = [("square", 12, 0, 1.0, 1.0, 0.5, 0.5, 3, 4, None), ("square", 12, 1, 1.0, 0.7, 0.0, 0.0, 3, 4, 0.3),
             ("honeycomb", 12, 0, 1.0, 1.0, 0.4, 0.4, 6, 2, 0.8), ("chain", 10, -1, 1.0, 1.0, 0.2, 0.2, 0, 0, 1.5),
             ("chain", 8, 0, 1.0, 1.0, 0.0, 0.0, 0, 0, 0.5), ("square", 8, 0, 1.0, 0.5, 0.3, 0.3, 2, 4, None)]
    lt = {"chain": O.CHAIN, "square": O.SQUARE, "honeycomb": O.HONEYCOMB}
    hams, refs = [], []
    for k, (lat, N, mag, J1, D1, J2, D2, lx, ly, W) in enumerate(specs):
        hi = None if W is None else Q.fields_from_seed(40 + k, N, -W, W)
        builds = ("build_diag", "build_off_diag") + (("build_disorder",) if hi is not None else ())
        hams.append(make(lat, N, mag, J1, D1, J2, D2, lx, ly, builds=builds, hi=hi))
        m = O.Model(lt[lat], N, N // 2 + mag, J1, D1, J2, D2, lx=lx, ly=ly)
        flags = 3 | (4 if hi is not None else 0)
        refs.append(np.linalg.eigvalsh(m.csr(hi, flags).toarray())[0])
    for kern in (1, 2):
        say("mixed batch, kernel", kern)
        e0 = BatchSolver(hams, tol=1e-12, hv_kernel=kern).solve()
        for k in range(len(hams)):
            assert abs(e0[k] - refs[k]) <= 1e-11 * abs(refs[k]), (kern, k, e0[k], refs[k])


def test_disorder_average_sqp_chunked(O, torch, tmp_path):
    """disorder_average with Sqp (S^+_i from Sz = 0 into Sz = 1, the target sector
    carrying the same fields) and batches smaller than the realisation count:
    energies and multiplet-summed weights of every file against the oracle."""
    import lsquantumed_amd as Q
    from lsquantumed_amd.disorder import disorder_average
    from test_oracle import _parse_dsf
    N, nev = 6, 5
    env = Q.Environment(N)
    lat = Q.chain1D(env)
    h0 = Q.Hamiltonian(env, Q.Basis(env, 0), lat, 1.0, 1.0, 0.0, 0.0)
    h1 = Q.Hamiltonian(env, Q.Basis(env, 1), lat, 1.0, 1.0, 0.0, 0.0)
    for h in (h0, h1):
        h.build_diag(env)
        h.build_off_diag(env)
    rd = Q.RandomDisorder(N, True, -0.5, 0.5, 3)
    done = disorder_average(env, h0, rd, str(tmp_path), nev=nev, op="Sqp", h1=h1, batch_size=2)
    assert sorted(done) == [0, 1, 2]
    coords = np.array([[i, 0] for i in range(N)], float)
    for k in range(3):
        hi = rd.hi_init(k).copy()
        w0, v0 = np.linalg.eigh(O.Model(O.CHAIN, N, 3, 1.0, 1.0, 0.0, 0.0).csr(hi, 7).toarray())
        w1, v1 = np.linalg.eigh(O.Model(O.CHAIN, N, 4, 1.0, 1.0, 0.0, 0.0).csr(hi, 7).toarray())
        gs, blocks = _parse_dsf(open(tmp_path / f"dsf_d{k}.dat").read())
        assert abs(gs - w0[0]) < 1e-11
        ex = np.array([e for e, _ in blocks[0]])
        np.testing.assert_allclose(ex, w1[:nev] - w0[0], rtol=0, atol=1e-10)
        amp = np.array([[blocks[i][n][1] for i in range(N)] for n in range(nev)])
        ref = O.dsf_amplitudes(N, 3, v0[:, 0], v1[:, :nev], "p")
        groups = O.multiplets(w1[:nev])
        if abs(w1[nev] - w1[nev - 1]) <= 1e-8:
            groups = groups[:-1]
        for grp in groups:
            for qx in range(N):
                q = (2 * np.pi * qx / N, 0.0)
                W = sum(O.dsf_weight(amp[n], q, coords) for n in grp)
                Wr = sum(O.dsf_weight(ref[n], q, coords) for n in grp)
                assert abs(W - Wr) < 1e-9, (k, grp, qx)
