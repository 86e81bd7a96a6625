"""lsqed_lanczos_step_noalloc (no device allocation; partial sums in a static
device array) against lsqed_lanczos_step through the C-ABI: identical alpha,
beta and vectors, bit for bit (same kernels, same reduction order), on two
calls in a row with a smaller sector on the second; LSQED_ERR_STATE when the
model's H.v plan has not been built.  Collected last (file name) so that a
failure here hides no other test."""
import ctypes as C
import sys

import pytest

pytestmark = pytest.mark.gpu

LSQED_ERR_STATE = -5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _model(Q, lat, N, lx=0, ly=0, J2=0.5, hi=None):
    env = Q.Environment(N)
    L = Q.square2D(env, lx, ly) if lat == "square" else Q.chain1D(env)
    h = Q.Hamiltonian(env, Q.Basis(env, 0), L, 1.0, 1.0, J2, J2)
    h.build_diag(env)
    h.build_off_diag(env)
    if hi is not None:
        h.build_disorder(env, hi)
    return h


def _step(fn, h, q, sq, p, sp, beta, y):
    from lsquantumed_amd._lib import ptr
    a, b = C.c_double(), C.c_double()
    rc = fn(h.hamilt, ptr(q), sq, ptr(p), sp, beta, ptr(y), C.byref(a), C.byref(b), None)
    return rc, a.value, b.value


def test_noalloc_step_matches_allocating_step():
    import numpy as np
    import torch

    import lsquantumed_amd as Q
    from lsquantumed_amd._lib import lib

    L = lib()
    cases = [("square", 24, 4, 6, None), ("chain", 16, 0, 0, np.linspace(-1, 1, 16))]  # second call: smaller
    for lat, N, lx, ly, hi in cases:
        say("noalloc step", lat, N)
        h = _model(Q, lat, N, lx, ly, hi=hi)
        D = h.hamilt_size()
        g = torch.Generator(device="cuda").manual_seed(N)
        q0 = torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5
        p0 = torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5
        if lat == "square":
            # before any H.v on this model its plan does not exist
            y = torch.empty_like(q0)
            rc, _, _ = _step(L.lsqed_lanczos_step_noalloc, h, q0, 0.5, p0.clone(), 0.25, 0.7, y)
            assert rc == LSQED_ERR_STATE
        results = []
        for fn in (L.lsqed_lanczos_step, L.lsqed_lanczos_step_noalloc, L.lsqed_lanczos_step_noalloc):
            p = p0.clone()
            y = torch.full_like(q0, float("nan"))
            rc, a, b = _step(fn, h, q0, 0.5, p, 0.25, 0.7, y)
            assert rc == 0, L.lsqed_last_error()
            torch.cuda.synchronize()
            results.append((a, b, p, y))
        ref = results[0]
        for a, b, p, y in results[1:]:
            assert a == ref[0] and b == ref[1], (a, ref[0], b, ref[1])
            assert torch.equal(p, ref[2]) and torch.equal(y, ref[3])
        say("  alpha", ref[0], "beta'", ref[1])
