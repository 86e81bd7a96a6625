"""Differentiable ground energy E0(h, J1, J2) of periodic chains with
longitudinal fields, over the batched solver of liblsqed.so.

    h  = torch.tensor([B, N], float64)        # fields, one row per realisation
    J1 = torch.tensor(1.0, requires_grad=True)
    E0 = ground_energy(h, J1, J2)             # [B]
    E0.sum().backward()

The model is Hamiltonian<chain1D>(J1, D1=J1, J2, D2=J2) with build_diag,
build_off_diag and build_disorder(h) (include/hamiltonian.tmpl_func.hpp
:69-344): H = sum_nn [J1 (S+S- + S-S+)/2 + J1 SzSz] + sum_nnn [J2 ... ]
+ sum_i h_i S^z_i in the Sz = 0 sector.  H is linear in (h, J1, J2), so by
Hellmann-Feynman
    dE0/dh_i = <S^z_i>,   dE0/dJ_k = sum_{bonds of kind k} <S_i . S_j>
in the ground state.  The forward pass keeps the ground states on the GPU;
the backward pass runs the batched observables kernel on them (one launch
for the whole batch) and contracts with the incoming gradient.
"""
from __future__ import annotations

import numpy as np

from .batch import OBS_FLIP_NN, OBS_FLIP_NNN, OBS_ZZ_NN, OBS_ZZ_NNN, BatchSolver
from .model import Basis, Environment, Hamiltonian, chain1D

_MODELS: dict = {}


def _models(B: int, N: int):
    """B chain models of N spins, reused across calls (couplings and fields
    are reset before every solve)."""
    key = (B, N)
    hams = _MODELS.get(key)
    if hams is None:
        env = Environment(N)
        basis = Basis(env, 0)
        lat = chain1D(env)
        hams = []
        for _ in range(B):
            h = Hamiltonian(env, basis, lat, 1.0, 1.0, 0.0, 0.0)
            h.build_diag(env)
            h.build_off_diag(env)
            hams.append(h)
        _MODELS[key] = hams
    return hams


def _set(hams, h: np.ndarray, J1: float, J2: float):
    for k, m in enumerate(hams):
        m.set_couplings(J1, J1, J2, J2)
        m.build_disorder(None, h[k])


def _torch_function():
    import torch

    class GroundEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, h, J1, J2, tol, maxit, seed):
            if h.dim() != 2:
                raise ValueError("h must be [batch, nspins]")
            B, N = h.shape
            hn = h.detach().to("cpu", torch.float64).numpy()
            j1, j2 = float(J1), float(J2)
            hams = _models(B, N)
            _set(hams, hn, j1, j2)
            need = any(ctx.needs_input_grad[:3])
            solver = BatchSolver(hams, maxit=maxit, tol=tol, seed=seed, keep_vectors=need)
            e0 = solver.solve()
            ctx.solver = solver if need else None
            ctx.N = N
            ctx.dev = h.device
            ctx.jdev = (J1.device if isinstance(J1, torch.Tensor) else h.device,
                        J2.device if isinstance(J2, torch.Tensor) else h.device)
            if not need:
                solver.destroy()
            return torch.as_tensor(e0, dtype=torch.float64, device=h.device)

        @staticmethod
        def backward(ctx, g):
            solver = ctx.solver
            if solver is None:
                return None, None, None, None, None, None
            obs = solver.observables()
            solver.destroy()
            ctx.solver = None
            N = ctx.N
            gn = g.detach().to("cpu", torch.float64).numpy()
            dh = obs[:, :N]
            dj1 = obs[:, OBS_ZZ_NN] + obs[:, OBS_FLIP_NN]
            dj2 = obs[:, OBS_ZZ_NNN] + obs[:, OBS_FLIP_NNN]
            gh = torch.as_tensor(gn[:, None] * dh, dtype=torch.float64, device=ctx.dev)
            gj1 = torch.as_tensor(float(gn @ dj1), dtype=torch.float64, device=ctx.jdev[0])
            gj2 = torch.as_tensor(float(gn @ dj2), dtype=torch.float64, device=ctx.jdev[1])
            return (gh if ctx.needs_input_grad[0] else None, gj1 if ctx.needs_input_grad[1] else None,
                    gj2 if ctx.needs_input_grad[2] else None, None, None, None)

    return GroundEnergy


_FN = None


def ground_energy(h, J1, J2, tol: float = 1e-12, maxit: int = 400, seed: int = 1):
    """E0 of each realisation (row of h): periodic chain of N = h.shape[1]
    spins, couplings J1 (nn) and J2 (nnn), isotropic, fields h, Sz = 0.
    Differentiable in h, J1 and J2 (0-dim tensors or floats)."""
    global _FN
    import torch

    if _FN is None:
        _FN = _torch_function()
    if not isinstance(J1, torch.Tensor):
        J1 = torch.tensor(float(J1), dtype=torch.float64)
    if not isinstance(J2, torch.Tensor):
        J2 = torch.tensor(float(J2), dtype=torch.float64)
    return _FN.apply(h, J1, J2, float(tol), int(maxit), int(seed))
