"""The complex-scalar boundary: lsqed_hv_c128 / lsqed_diag_c128 on
interleaved complex128 vectors, the layout of the complex PETSc the reference
requires (README.md:18), against the oracle's complex128 CSR SpMV (PETSc's
MATMPIAIJ layout: complex values, int64 columns) and, part by part, against
the real product (bit for bit: H is real and each part runs the real
kernel's arithmetic)."""
import sys

import numpy as np
import pytest

from test_gpu_hamiltonian import make, oracle_model

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available()
    return T


def _cvec(torch, D, seed):
    g = np.random.default_rng(seed)
    x = g.uniform(-1, 1, D) + 1j * g.uniform(-1, 1, D)
    return x, torch.from_numpy(x).cuda()


@pytest.mark.parametrize("lat,N,mag,lx,ly,J2,fields", [
    ("square", 24, 0, 4, 6, 0.5, False), ("chain", 20, 1, 0, 0, 0.2, True), ("honeycomb", 16, -1, 4, 4, 0.4, True),
    ("square", 12, 0, 3, 4, 1.0, True)])
@pytest.mark.parametrize("kernel", [1, 3], ids=["generic", "split_jit"])
def test_hv_c128_vs_oracle_complex_csr(O, torch, lat, N, mag, lx, ly, J2, fields, kernel):
    say("hv c128", lat, N, mag, fields, kernel)
    hi = np.random.default_rng(N).uniform(-2, 2, N) if fields else None
    builds = ("build_diag", "build_off_diag") + (("build_disorder",) if fields else ())
    h = make(lat, N, mag, 1.0, 1.0, J2, J2, lx, ly, builds=builds, hi=hi)
    h.set_hv_kernel(kernel)
    m = oracle_model(O, lat, N, mag, 1.0, 1.0, J2, J2, lx, ly)
    A = O.ComplexCsr(m, hi, 7 if fields else 3)
    D = h.hamilt_size()
    xh, x = _cvec(torch, D, 7)
    y = h.matvec_c128(x)
    yref = np.empty_like(xh)
    A.matvec(np.ascontiguousarray(xh), yref)
    err = np.max(np.abs(y.cpu().numpy() - yref))
    assert err <= 1e-13 * np.max(np.abs(yref)), err
    # each part equals the real product of that part, bit for bit
    yr = h.matvec(x.real.contiguous())
    yi = h.matvec(x.imag.contiguous())
    assert torch.equal(y.real, yr) and torch.equal(y.imag, yi)
    # the complex diagonal: real parts = lsqed_diag, imaginary parts 0
    dc = h.diagonal(complex128=True)
    assert torch.equal(dc.real, h.diagonal()) and not torch.any(dc.imag != 0)


def test_hv_c128_4x8_rows(O, torch):
    """4x8 J1-J2 (D = 601,080,390; 9.6 GB per complex vector): rows across
    the sector against the oracle's matrix-free product of each part, the
    parts against the real product bit for bit, and <x, H z> = <H x, z>."""
    say("hv c128 4x8")
    h = make("square", 32, 0, 1.0, 1.0, 0.5, 0.5, 4, 8)
    assert h.hv_kernel() == 3
    m = oracle_model(O, "square", 32, 0, 1.0, 1.0, 0.5, 0.5, 4, 8)
    D = h.hamilt_size()
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.complex(torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5,
                      torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5)
    y = h.matvec_c128(x)
    torch.cuda.synchronize()
    for part in ("real", "imag"):
        xp = getattr(x, part).contiguous()
        yp = getattr(y, part).contiguous()
        assert torch.equal(yp, h.matvec(xp)), part
        xh = xp.cpu().numpy()
        yh = yp.cpu().numpy()
        for row0 in (0, 98765432, D // 2, D - 4000):
            yref = m.hv_matfree(xh, row0=row0, nrows=4000)
            assert np.max(np.abs(yh[row0:row0 + 4000] - yref)) <= 1e-12 * np.abs(yref).max(), (part, row0)
        del xh, yh, xp, yp
    z = torch.complex(torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5,
                      torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5)
    hz = h.matvec_c128(z)
    a = torch.vdot(x, hz).item()
    b = torch.vdot(y, z).item()
    assert abs(a - b) <= 1e-10 * (abs(a) + abs(b)), (a, b)


@pytest.mark.parametrize("kernel", [1, 3], ids=["generic", "split_jit"])
def test_hv_c128_36_sites_wide_ranks(O, torch, kernel):
    """The complex product on a 36-site lattice (64-bit states and row bases:
    the wide forms of both kernels), 6x6 J1-J2 in the sector with 10 up spins
    (D = 254,186,856): x = (1 + 2i) f with f the oracle's formula vector of
    the rank; rows across the sector equal (1 + 2i) times the oracle's rows,
    and each part equals the real product of that part bit for bit."""
    from test_gpu_hamiltonian import formula_vector
    say("hv c128 36 sites", kernel)
    h = make("square", 36, 10 - 18, 1.0, 1.0, 0.5, 0.5, 6, 6)
    h.set_hv_kernel(kernel)
    D = h.hamilt_size()
    assert D == 254186856
    f = formula_vector(O, torch, D)
    x = torch.complex(f, 2.0 * f)
    y = h.matvec_c128(x)
    assert torch.equal(y.real, h.matvec(x.real.contiguous()))
    assert torch.equal(y.imag, h.matvec(x.imag.contiguous()))
    m = oracle_model(O, "square", 36, 10 - 18, 1.0, 1.0, 0.5, 0.5, 6, 6)
    for row0 in (0, D // 3, D - 3000):
        yref = m.hv_rows_formula(row0, 3000)
        got = y[row0:row0 + 3000].cpu().numpy()
        scale = np.abs(yref).max()
        assert np.max(np.abs(got.real - yref)) <= 1e-12 * scale, row0
        assert np.max(np.abs(got.imag - 2.0 * yref)) <= 2e-12 * scale, row0
    del f, x, y
    torch.cuda.empty_cache()
