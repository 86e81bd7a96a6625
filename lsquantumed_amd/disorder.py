"""Disorder average on the device: the realisation loop of
examples/main_disorder.cpp:82-120.

The reference, per realisation k: RandomDisorder::hi_init(k), build_disorder,
a Krylov-Schur solve for the ground state (nev = 1), then
Phys::DynStructFactor, which solves the nev lowest states of the same
sector and writes path/dsf_d<k>.dat.  Here the ground states of all pending
realisations are solved together by the batched Lanczos (one model per
realisation, ground-state vectors kept), and each realisation's DSF is then
computed as in the reference.  With resume=True a realisation whose
dsf_d<k>.dat already exists is skipped, so an interrupted run continues where
it stopped (the files are the only state the reference persists).
"""
from __future__ import annotations

import os

from .batch import BatchSolver
from .model import Hamiltonian
from .phys import Phys


def dsf_path(path: str, k: int) -> str:
    return os.path.join(path, f"dsf_d{k}.dat")


def disorder_average(env, h, rd, path: str, nev: int = 10, op: str = "Sqz", h1=None, resume: bool = True,
                     tol: float = 1e-12, dsf_tol: float = 1e-12, batch_size: int = 256):
    """Runs every realisation of `rd` (Tools::RandomDisorder) on the model of
    `h` (lattice, couplings and build calls; its fields are replaced per
    realisation) and writes dsf_d<k>.dat like main_disorder.cpp.  `h1` is the
    target-sector model for Sqp (None: the same sector, Sqz).  Returns
    {k: E0} for the realisations computed in this call."""
    if op not in ("Sqz", "Sqp"):
        raise ValueError("op must be Sqz or Sqp")
    os.makedirs(path, exist_ok=True)
    pending = [k for k in range(rd.rep) if not (resume and os.path.exists(dsf_path(path, k)))]
    done = {}
    for c0 in range(0, len(pending), batch_size):
        chunk = pending[c0:c0 + batch_size]
        models, fields = [], []
        for k in chunk:
            hi = rd.hi_init(k).copy()
            m = Hamiltonian(env, h.basis, h.lattice, h.J1, h.Delta1, h.J2, h.Delta2)
            if h._flags & 1:
                m.build_diag(env)
            if h._flags & 2:
                m.build_off_diag(env)
            m.build_disorder(env, hi)
            models.append(m)
            fields.append(hi)
        b = BatchSolver(models, tol=tol, keep_vectors=True)
        e0 = b.solve()
        for j, k in enumerate(chunk):
            psi0 = b.ground_state(j)
            target = models[j]
            if op == "Sqp":
                if h1 is None:
                    raise ValueError("Sqp needs the target-sector model h1")
                target = Hamiltonian(env, h1.basis, h1.lattice, h1.J1, h1.Delta1, h1.J2, h1.Delta2)
                if h1._flags & 1:
                    target.build_diag(env)
                if h1._flags & 2:
                    target.build_off_diag(env)
                target.build_disorder(env, fields[j])
            data = dict(h0=models[j], h1=target, nev=nev, ncv=-2, mpd=-2, disorder=True, hi=fields[j], path=path,
                        tol=dsf_tol)
            Phys.DynStructFactor(env, data, float(e0[j]), psi0, k, op=op)
            done[k] = float(e0[j])
        b.destroy()
    return done
